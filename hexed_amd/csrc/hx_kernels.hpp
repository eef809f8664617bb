// gfx950 kernels of the DG hot path, FP64 form. Included once per n_dim by hx_kernels_nd*.hip, which
// instantiate them for row sizes 2..8.
//
// Element kernels run one thread per quadrature point and EPB = max(1, 256/nq) elements per workgroup; the
// per-element flux tensor, face values and new state are staged in LDS so that every derivative contraction
// and face extrapolation reads on-chip data. Connection kernels run one thread per face quadrature point.
#pragma once

#include <cfloat>

#include "hx_launch.hpp"

namespace hx {

template <int ND, int RS>
struct Shape
{
  static constexpr int NQ = ipow(RS, ND);
  static constexpr int NFQ = NQ / RS;
  static constexpr int NV = ND + 2;
  static constexpr int NVERT = 1 << ND;
  static constexpr int EPB = (256 / NQ) > 0 ? 256 / NQ : 1;
  static constexpr int BLOCK = EPB * NQ;
  static constexpr int NFACE_DOF = 2 * ND * NV * NFQ;  // face values per element
};

template <int RS>
__device__ inline void load_ops(Ops<RS>& s, const Ops<RS>* g, int t, int nthreads)
{
  const double* src = reinterpret_cast<const double*>(g);
  double* dst = reinterpret_cast<double*>(&s);
  for (int i = t; i < int(sizeof(Ops<RS>) / sizeof(double)); i += nthreads) dst[i] = src[i];
}

__device__ inline double* slot_ptr(const Dev_mesh& m, int slot) { return m.faces + size_t(slot) * m.face_sz; }

// ---- face extrapolation of element data held in LDS (Spatial.hpp:41-57) ----------------------------------
// src: [EPB][NV][NQ] in LDS; writes block `ldg` (0: state/flux, 1: LDG) of every face slot.
template <int ND, int RS>
__device__ inline void extrapolate_faces(const Dev_mesh& m, int cls, const Ops<RS>& ops, const double* src, int ldg)
{
  using S = Shape<ND, RS>;
  using R = Row<ND, RS>;
  const int n = m.n_elem[cls];
  for (int k = threadIdx.x; k < S::EPB * S::NFACE_DOF; k += S::BLOCK) {
    const int el = k / S::NFACE_DOF;
    const int r = k % S::NFACE_DOF;
    const int f = r / (S::NV * S::NFQ);
    const int v = (r / S::NFQ) % S::NV;
    const int fq = r % S::NFQ;
    const int e = blockIdx.x * S::EPB + el;
    if (e >= n) continue;
    const int i = f / 2, s = f % 2;
    const int b = R::base_of_fq(fq, i), sd = R::stride(i);
    const double* row = src + (el * S::NV + v) * S::NQ + b;
    double acc = 0;
    for (int kk = 0; kk < RS; ++kk) acc += ops.bound[s][kk] * row[kk * sd];
    slot_ptr(m, m.face_slot[cls][e * 2 * ND + f])[ldg * S::NV * S::NFQ + v * S::NFQ + fq] = acc;
  }
}

// Modal filter along every dimension in turn (Spatial.hpp:472-481): buf [EPB][nvar][NQ] in LDS, vals per thread.
template <int ND, int RS, int NVAR>
__device__ inline void filter_rows(const Ops<RS>& ops, double* buf, int el, int q, double (&vals)[NVAR])
{
  using S = Shape<ND, RS>;
  using R = Row<ND, RS>;
  double* mine = buf + el * NVAR * S::NQ;
  __syncthreads();
  for (int v = 0; v < NVAR; ++v) mine[v * S::NQ + q] = vals[v];
  for (int i = 0; i < ND; ++i) {
    __syncthreads();
    const int ki = R::node(q, i), b = R::base(q, i), sd = R::stride(i);
    for (int v = 0; v < NVAR; ++v) {
      double acc = 0;
      for (int k = 0; k < RS; ++k) acc += ops.filt[ki][k] * mine[v * S::NQ + b + k * sd];
      vals[v] = acc;
    }
    __syncthreads();
    for (int v = 0; v < NVAR; ++v) mine[v * S::NQ + q] = vals[v];
  }
  __syncthreads();
}

// ---- Local, Euler (Spatial.hpp:326-509 with Navier_stokes<false>) -----------------------------------------
template <int ND, int RS, bool DEF>
__global__ __launch_bounds__((Shape<ND, RS>::BLOCK)) void k_local_euler(Dev_mesh m, const Ops<RS>* __restrict__ gops,
                                                                      double coef, int stage, int compute_residual,
                                                                      int use_filter)
{
  using S = Shape<ND, RS>;
  using R = Row<ND, RS>;
  constexpr int NQ = S::NQ, NFQ = S::NFQ, NV = S::NV, EPB = S::EPB;
  constexpr int cls = DEF ? HX_DEFORMED : HX_CARTESIAN;
  __shared__ Ops<RS> ops;
  __shared__ double s_flux[EPB][ND][NV][NQ];
  __shared__ double s_face[EPB][2 * ND][NV][NFQ];
  const int t = threadIdx.x, el = t / NQ, q = t % NQ;
  const int n = m.n_elem[cls];
  const int e = blockIdx.x * EPB + el;
  const bool valid = e < n;
  load_ops(ops, gops, t, S::BLOCK);

  double u[NV];
  const double* st = m.state[cls] + size_t(e) * NV * NQ;
  for (int v = 0; v < NV; ++v) u[v] = valid ? st[v * NQ + q] : (v == ND ? 1. : 0.);
  double det = 1.;
  double met[ND][ND];
  if constexpr (DEF) {
    const double* mt = m.metrics + size_t(e) * (ND * ND + 1) * NQ;
    for (int i = 0; i < ND; ++i)
      for (int j = 0; j < ND; ++j) met[i][j] = valid ? mt[(i * ND + j) * NQ + q] : double(i == j);
    det = valid ? mt[ND * ND * NQ + q] : 1.;
  }
  // volume flux along each reference direction (Spatial.hpp:404-419)
  for (int i = 0; i < ND; ++i) {
    double nrm[ND];
    for (int j = 0; j < ND; ++j) {
      if constexpr (DEF) nrm[j] = met[i][j];
      else nrm[j] = (i == j);
    }
    double f[NV], p;
    flux_conv<ND>(u, nrm, f, p);
    for (int v = 0; v < NV; ++v) s_flux[el][i][v][q] = f[v];
  }
  // numerical face fluxes written by Neighbor
  for (int k = t; k < EPB * S::NFACE_DOF; k += S::BLOCK) {
    const int el2 = k / S::NFACE_DOF, r = k % S::NFACE_DOF, f = r / (NV * NFQ), vq = r % (NV * NFQ);
    const int e2 = blockIdx.x * EPB + el2;
    (&s_face[el2][f][0][0])[vq] = e2 < n ? slot_ptr(m, m.face_slot[cls][e2 * 2 * ND + f])[vq] : 0.;
  }
  __syncthreads();
  // -derivative(row_f, face_f) along every dimension (Spatial.hpp:444-456, Derivative.hpp:51-55)
  double rate[NV] = {};
  for (int i = 0; i < ND; ++i) {
    const int ki = R::node(q, i), b = R::base(q, i), sd = R::stride(i), fq = R::fq(q, i);
    for (int v = 0; v < NV; ++v) {
      double dq = 0, b0 = 0, b1 = 0;
      for (int k = 0; k < RS; ++k) {
        const double fk = s_flux[el][i][v][b + k * sd];
        dq += ops.diff[ki][k] * fk;
        b0 += ops.bound[0][k] * fk;
        b1 += ops.bound[1][k] * fk;
      }
      const double j0 = s_face[el][2 * i][v][fq] - b0;
      const double j1 = s_face[el][2 * i + 1][v][fq] - b1;
      rate[v] -= dq + (ops.lift[ki][0] * j0 + ops.lift[ki][1] * j1);
    }
  }
  if (use_filter) filter_rows<ND, RS, NV>(ops, &s_flux[0][0][0][0], el, q, rate);
  // stage update (Spatial.hpp:483-503)
  if (valid) {
    double mult = coef * m.tss[cls][size_t(e) * NQ + q] / m.nom_size[cls][e];
    if constexpr (DEF) mult /= det;
    double* rc = m.rcache[cls] + size_t(e) * NV * NQ;
    double* sw = m.state[cls] + size_t(e) * NV * NQ;
    for (int v = 0; v < NV; ++v) {
      double r = rate[v];
      if (stage) r -= rc[v * NQ + q];
      else rc[v * NQ + q] = r;
      r *= mult;
      if (compute_residual) rc[v * NQ + q] = r;
      else u[v] += r;
    }
    if (!compute_residual)
      for (int v = 0; v < NV; ++v) sw[v * NQ + q] = u[v];
  }
  // Write_face of the updated state (Spatial.hpp:507)
  __syncthreads();
  for (int v = 0; v < NV; ++v) s_flux[el][0][v][q] = u[v];
  __syncthreads();
  double* su = &s_flux[0][0][0][0];
  // s_flux[el] is ND*NV*NQ long; the state block of element el sits at its start
  using Sh = S;
  for (int k = t; k < EPB * Sh::NFACE_DOF; k += Sh::BLOCK) {
    const int el2 = k / Sh::NFACE_DOF, r = k % Sh::NFACE_DOF;
    const int f = r / (NV * NFQ), v = (r / NFQ) % NV, fq = r % NFQ;
    const int e2 = blockIdx.x * EPB + el2;
    if (e2 >= n) continue;
    const int i = f / 2, s = f % 2;
    const int b = R::base_of_fq(fq, i), sd = R::stride(i);
    const double* row = su + (el2 * ND * NV + v) * NQ + b;
    double acc = 0;
    for (int kk = 0; kk < RS; ++kk) acc += ops.bound[s][kk] * row[kk * sd];
    slot_ptr(m, m.face_slot[cls][e2 * 2 * ND + f])[v * NFQ + fq] = acc;
  }
}

// ---- Local, Navier-Stokes stage 0 (Spatial.hpp:326-509 with Navier_stokes<true>) -------------------------
template <int ND, int RS, bool DEF>
struct Ns_lds
{
  using S = Shape<ND, RS>;
  // phase 1: gradient inputs; phase 2: fluxes. Both start at offset 0 of the element's region.
  static constexpr int P1 = S::NV * S::NQ + 2 * ND * S::NV * S::NFQ + (DEF ? ND * ND * S::NQ + 2 * ND * ND * S::NFQ : 0);
  static constexpr int P2 = 2 * ND * S::NV * S::NQ + 2 * ND * S::NV * S::NFQ;
  static constexpr int PER_ELEM = P1 > P2 ? P1 : P2;
};

template <int ND, int RS, bool DEF>
__global__ __launch_bounds__((Shape<ND, RS>::BLOCK)) void k_local_ns(Dev_mesh m, const Ops<RS>* __restrict__ gops,
                                                                   Ns_params ns, double coef, int compute_residual,
                                                                   int use_filter)
{
  using S = Shape<ND, RS>;
  using R = Row<ND, RS>;
  using L = Ns_lds<ND, RS, DEF>;
  constexpr int NQ = S::NQ, NFQ = S::NFQ, NV = S::NV, EPB = S::EPB;
  constexpr int cls = DEF ? HX_DEFORMED : HX_CARTESIAN;
  __shared__ Ops<RS> ops;
  __shared__ double lds[EPB * L::PER_ELEM];
  const int t = threadIdx.x, el = t / NQ, q = t % NQ;
  const int n = m.n_elem[cls];
  const int e = blockIdx.x * EPB + el;
  const bool valid = e < n;
  load_ops(ops, gops, t, S::BLOCK);
  double* my = lds + el * L::PER_ELEM;
  // phase 1 layout
  double* su = my;                                  // [NV][NQ]
  double* sl = su + NV * NQ;                        // [2ND][NV][NFQ] LDG face state
  double* sn = sl + 2 * ND * NV * NFQ;              // [ND][ND][NQ] reference-level normals (DEF)
  double* sfn = sn + ND * ND * NQ;                  // [2ND][ND][NFQ] face normals (DEF)

  double u[NV];
  const double* st = m.state[cls] + size_t(e) * NV * NQ;
  for (int v = 0; v < NV; ++v) {
    u[v] = valid ? st[v * NQ + q] : (v == ND ? 1. : (v == ND + 1 ? 1. : 0.));
    su[v * NQ + q] = u[v];
  }
  double bulk = 0., lap = 0.;
  if (valid && m.av[cls]) {
    bulk = m.av[cls][size_t(e) * 2 * NQ + q];
    lap = m.av[cls][size_t(e) * 2 * NQ + NQ + q];
  }
  double det = 1.;
  double met[ND][ND];
  if constexpr (DEF) {
    const double* mt = m.metrics + size_t(e) * (ND * ND + 1) * NQ;
    for (int i = 0; i < ND; ++i)
      for (int j = 0; j < ND; ++j) {
        met[i][j] = valid ? mt[(i * ND + j) * NQ + q] : double(i == j);
        sn[(i * ND + j) * NQ + q] = met[i][j];
      }
    det = valid ? mt[ND * ND * NQ + q] : 1.;
  }
  // LDG face states (and face normals) of every element in the block
  for (int k = t; k < EPB * 2 * ND * NV * NFQ; k += S::BLOCK) {
    const int el2 = k / (2 * ND * NV * NFQ), r = k % (2 * ND * NV * NFQ), f = r / (NV * NFQ), vq = r % (NV * NFQ);
    const int e2 = blockIdx.x * EPB + el2;
    lds[el2 * L::PER_ELEM + NV * NQ + r] = e2 < n ? slot_ptr(m, m.face_slot[cls][e2 * 2 * ND + f])[NV * NFQ + vq] : 0.;
  }
  if constexpr (DEF) {
    for (int k = t; k < EPB * 2 * ND * ND * NFQ; k += S::BLOCK) {
      const int el2 = k / (2 * ND * ND * NFQ), r = k % (2 * ND * ND * NFQ), f = r / (ND * NFQ), jq = r % (ND * NFQ);
      const int e2 = blockIdx.x * EPB + el2;
      double val = 0.;
      if (e2 < n) {
        const int ns_ = m.face_nslot[e2 * 2 * ND + f];
        // a face in a Cartesian connection uses the identity normal (Spatial.hpp:331-339,364-367)
        val = ns_ >= 0 ? m.normals[size_t(ns_) * ND * NFQ + jq] : double(jq / NFQ == f / 2);
      }
      lds[el2 * L::PER_ELEM + NV * NQ + 2 * ND * NV * NFQ + ND * ND * NQ + r] = val;
    }
  }
  __syncthreads();
  // gradient times Jacobian determinant (Spatial.hpp:371-402), divided by nominal size
  const double h = valid ? m.nom_size[cls][e] : 1.;
  double grad[NV][ND];
  for (int v = 0; v < NV; ++v)
    for (int j = 0; j < ND; ++j) grad[v][j] = 0.;
  for (int i = 0; i < ND; ++i) {
    const int ki = R::node(q, i), b = R::base(q, i), sd = R::stride(i), fq = R::fq(q, i);
    if constexpr (DEF) {
      for (int j = 0; j < ND; ++j) {
        double nk[RS];
        for (int k = 0; k < RS; ++k) nk[k] = sn[(i * ND + j) * NQ + b + k * sd];
        const double fn0 = sfn[((2 * i) * ND + j) * NFQ + fq], fn1 = sfn[((2 * i + 1) * ND + j) * NFQ + fq];
        for (int v = 0; v < NV; ++v) {
          double dq = 0, b0 = 0, b1 = 0;
          for (int k = 0; k < RS; ++k) {
            const double x = nk[k] * su[v * NQ + b + k * sd];
            dq += ops.diff[ki][k] * x;
            b0 += ops.bound[0][k] * x;
            b1 += ops.bound[1][k] * x;
          }
          const double j0 = fn0 * sl[((2 * i) * NV + v) * NFQ + fq] - b0;
          const double j1 = fn1 * sl[((2 * i + 1) * NV + v) * NFQ + fq] - b1;
          grad[v][j] += (dq + (ops.lift[ki][0] * j0 + ops.lift[ki][1] * j1)) / h;
        }
      }
    } else {
      for (int v = 0; v < NV; ++v) {
        double dq = 0, b0 = 0, b1 = 0;
        for (int k = 0; k < RS; ++k) {
          const double x = su[v * NQ + b + k * sd];
          dq += ops.diff[ki][k] * x;
          b0 += ops.bound[0][k] * x;
          b1 += ops.bound[1][k] * x;
        }
        const double j0 = sl[((2 * i) * NV + v) * NFQ + fq] - b0;
        const double j1 = sl[((2 * i + 1) * NV + v) * NFQ + fq] - b1;
        grad[v][i] = (dq + (ops.lift[ki][0] * j0 + ops.lift[ki][1] * j1)) / h;
      }
    }
  }
  // volume fluxes (Spatial.hpp:404-435)
  double fc[ND][NV];
  double fv[NV][ND];
  {
    double nrm_t[ND][ND];  // normal(j, i) = met[i][j]
    for (int i = 0; i < ND; ++i) {
      double nrm[ND];
      for (int j = 0; j < ND; ++j) {
        if constexpr (DEF) nrm[j] = met[i][j];
        else nrm[j] = (i == j);
        nrm_t[j][i] = nrm[j];
      }
      double f[NV], p;
      flux_conv<ND>(u, nrm, f, p);
      for (int v = 0; v < NV; ++v) fc[i][v] = f[v];
    }
    if constexpr (DEF)
      for (int v = 0; v < NV; ++v)
        for (int j = 0; j < ND; ++j) grad[v][j] /= det;
    flux_diff<ND, DEF>(ns, u, bulk, lap, grad, nrm_t, fv);
  }
  __syncthreads();  // phase 1 data is dead from here on
  double* sf = my;                                  // [ND][NV][NQ] convective flux
  double* sv = sf + ND * NV * NQ;                   // [ND][NV][NQ] viscous flux
  double* sb = sv + ND * NV * NQ;                   // [2ND][NV][NFQ] numerical face flux
  for (int i = 0; i < ND; ++i)
    for (int v = 0; v < NV; ++v) {
      sf[(i * NV + v) * NQ + q] = fc[i][v];
      sv[(i * NV + v) * NQ + q] = fv[v][i];
    }
  for (int k = t; k < EPB * 2 * ND * NV * NFQ; k += S::BLOCK) {
    const int el2 = k / (2 * ND * NV * NFQ), r = k % (2 * ND * NV * NFQ), f = r / (NV * NFQ), vq = r % (NV * NFQ);
    const int e2 = blockIdx.x * EPB + el2;
    lds[el2 * L::PER_ELEM + 2 * ND * NV * NQ + r] = e2 < n ? slot_ptr(m, m.face_slot[cls][e2 * 2 * ND + f])[vq] : 0.;
  }
  __syncthreads();
  // residual (Spatial.hpp:444-470)
  double rate[2 * NV] = {};
  for (int i = 0; i < ND; ++i) {
    const int ki = R::node(q, i), b = R::base(q, i), sd = R::stride(i), fq = R::fq(q, i);
    for (int v = 0; v < NV; ++v) {
      double dq = 0, b0 = 0, b1 = 0, dv = 0;
      for (int k = 0; k < RS; ++k) {
        const double fk = sf[(i * NV + v) * NQ + b + k * sd];
        dq += ops.diff[ki][k] * fk;
        b0 += ops.bound[0][k] * fk;
        b1 += ops.bound[1][k] * fk;
        dv += ops.diff[ki][k] * sv[(i * NV + v) * NQ + b + k * sd];
      }
      const double j0 = sb[((2 * i) * NV + v) * NFQ + fq] - b0;
      const double j1 = sb[((2 * i + 1) * NV + v) * NFQ + fq] - b1;
      rate[v] -= dq + (ops.lift[ki][0] * j0 + ops.lift[ki][1] * j1);
      rate[NV + v] -= dv;
    }
  }
  // viscous flux extrapolated to the LDG face blocks (Spatial.hpp:460-466)
  for (int k = t; k < EPB * S::NFACE_DOF; k += S::BLOCK) {
    const int el2 = k / S::NFACE_DOF, r = k % S::NFACE_DOF;
    const int f = r / (NV * NFQ), v = (r / NFQ) % NV, fq = r % NFQ;
    const int e2 = blockIdx.x * EPB + el2;
    if (e2 >= n) continue;
    const int i = f / 2, s = f % 2;
    const int b = R::base_of_fq(fq, i), sd = R::stride(i);
    const double* row = lds + el2 * L::PER_ELEM + ND * NV * NQ + (i * NV + v) * NQ + b;
    double acc = 0;
    for (int kk = 0; kk < RS; ++kk) acc += ops.bound[s][kk] * row[kk * sd];
    slot_ptr(m, m.face_slot[cls][e2 * 2 * ND + f])[NV * NFQ + v * NFQ + fq] = acc;
  }
  if (use_filter) filter_rows<ND, RS, 2 * NV>(ops, lds, el, q, rate);
  // stage-0 update (Spatial.hpp:483-503): the convective part goes to the residual cache
  if (valid) {
    double mult = coef * m.tss[cls][size_t(e) * NQ + q] / h;
    if constexpr (DEF) mult /= det;
    double* rc = m.rcache[cls] + size_t(e) * NV * NQ;
    double* sw = m.state[cls] + size_t(e) * NV * NQ;
    for (int v = 0; v < NV; ++v) {
      double r = rate[v];
      rc[v * NQ + q] = r;
      r += rate[NV + v];
      r *= mult;
      if (compute_residual) rc[v * NQ + q] = r;
      else sw[v * NQ + q] = u[v] + r;
    }
  }
}

// ---- Reconcile_ldg_flux (Spatial.hpp:543-594) ------------------------------------------------------------
template <int ND, int RS, bool DEF>
__global__ __launch_bounds__((Shape<ND, RS>::BLOCK)) void k_reconcile_ldg(Dev_mesh m, const Ops<RS>* __restrict__ gops,
                                                                        double dt, int compute_residual, int use_filter)
{
  using S = Shape<ND, RS>;
  using R = Row<ND, RS>;
  constexpr int NQ = S::NQ, NFQ = S::NFQ, NV = S::NV, EPB = S::EPB;
  constexpr int cls = DEF ? HX_DEFORMED : HX_CARTESIAN;
  __shared__ Ops<RS> ops;
  __shared__ double su[EPB][NV][NQ];
  const int t = threadIdx.x, el = t / NQ, q = t % NQ;
  const int n = m.n_elem[cls];
  const int e = blockIdx.x * EPB + el;
  const bool valid = e < n;
  load_ops(ops, gops, t, S::BLOCK);
  double rate[NV] = {};
  if (valid) {
    for (int i = 0; i < ND; ++i) {
      const int ki = R::node(q, i), fq = R::fq(q, i);
      const double* f0 = slot_ptr(m, m.face_slot[cls][e * 2 * ND + 2 * i]) + NV * NFQ;
      const double* f1 = slot_ptr(m, m.face_slot[cls][e * 2 * ND + 2 * i + 1]) + NV * NFQ;
      for (int v = 0; v < NV; ++v)
        rate[v] -= ops.lift[ki][0] * f0[v * NFQ + fq] + ops.lift[ki][1] * f1[v * NFQ + fq];
    }
  }
  if (use_filter) filter_rows<ND, RS, NV>(ops, &su[0][0][0], el, q, rate);
  double u[NV];
  const double* st = m.state[cls] + size_t(e) * NV * NQ;
  for (int v = 0; v < NV; ++v) u[v] = valid ? st[v * NQ + q] : 0.;
  if (valid) {
    double mult = dt * m.tss[cls][size_t(e) * NQ + q] / m.nom_size[cls][e];
    if constexpr (DEF) mult /= m.metrics[size_t(e) * (ND * ND + 1) * NQ + ND * ND * NQ + q];
    if (compute_residual) {
      double* rc = m.rcache[cls] + size_t(e) * NV * NQ;
      for (int v = 0; v < NV; ++v) rc[v * NQ + q] += rate[v] * mult;
    } else {
      double* sw = m.state[cls] + size_t(e) * NV * NQ;
      for (int v = 0; v < NV; ++v) {
        u[v] += rate[v] * mult;
        sw[v * NQ + q] = u[v];
      }
    }
  }
  __syncthreads();
  for (int v = 0; v < NV; ++v) su[el][v][q] = u[v];
  __syncthreads();
  extrapolate_faces<ND, RS>(m, cls, ops, &su[0][0][0], 0);
}

// ---- Write_face (Spatial.hpp:60-70) ----------------------------------------------------------------------
template <int ND, int RS>
__global__ __launch_bounds__((Shape<ND, RS>::BLOCK)) void k_write_face(Dev_mesh m, const Ops<RS>* __restrict__ gops, int cls)
{
  using S = Shape<ND, RS>;
  constexpr int NQ = S::NQ, NV = S::NV, EPB = S::EPB;
  __shared__ Ops<RS> ops;
  __shared__ double su[EPB][NV][NQ];
  const int t = threadIdx.x, el = t / NQ, q = t % NQ;
  const int e = blockIdx.x * EPB + el;
  load_ops(ops, gops, t, S::BLOCK);
  const double* st = m.state[cls] + size_t(e) * NV * NQ;
  for (int v = 0; v < NV; ++v) su[el][v][q] = e < m.n_elem[cls] ? st[v * NQ + q] : 0.;
  __syncthreads();
  extrapolate_faces<ND, RS>(m, cls, ops, &su[0][0][0], 0);
}

// ---- Neighbor (Spatial.hpp:613-704) ----------------------------------------------------------------------
template <int ND, int RS, bool DEF, bool VISC>
__global__ __launch_bounds__(256) void k_neighbor(Dev_mesh m)
{
  constexpr int NFQ = Shape<ND, RS>::NFQ, NV = ND + 2;
  constexpr int cls = DEF ? HX_DEFORMED : HX_CARTESIAN;
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int con = gid / NFQ, f = gid % NFQ;
  if (con >= m.n_con[cls]) return;
  const Dir dir(m.con_dir[cls][con]);
  double* f0 = slot_ptr(m, m.con_slot[cls][2 * con]);
  double* f1 = slot_ptr(m, m.con_slot[cls][2 * con + 1]);
  int p1 = f;
  double sg0 = 1., sg1 = 1.;
  double n[ND];
  if constexpr (DEF) {
    p1 = face1_index<ND, RS>(dir, f);
    sg0 = 1 - 2 * int(dir.flip_normal(0));
    sg1 = 1 - 2 * int(dir.flip_normal(1));
    const double* nn = m.normals + size_t(m.con_nslot[2 * con]) * ND * NFQ;
    for (int j = 0; j < ND; ++j) n[j] = sg0 * nn[j * NFQ + f];
  } else {
    for (int j = 0; j < ND; ++j) n[j] = (j == dir.i_dim0);
  }
  double u0[NV], u1[NV];
  for (int v = 0; v < NV; ++v) {
    u0[v] = f0[v * NFQ + f];
    u1[v] = f1[v * NFQ + p1];
  }
  if constexpr (VISC) {
    for (int v = 0; v < NV; ++v) {
      const double avg = .5 * (u0[v] + u1[v]);
      f0[NV * NFQ + v * NFQ + f] = avg;
      f1[NV * NFQ + v * NFQ + p1] = avg;
    }
  }
  double F0[NV], F1[NV], p;
  flux_conv<ND>(u0, n, F0, p);
  flux_conv<ND>(u1, n, F1, p);
  double nsq = 0;
  for (int j = 0; j < ND; ++j) nsq += n[j] * n[j];
  const double c = fmax(char_speed<ND>(u0), char_speed<ND>(u1)) * sqrt(nsq);
  for (int v = 0; v < NV; ++v) {
    const double fl = .5 * (F0[v] + F1[v] + c * (u0[v] - u1[v]));
    f0[v * NFQ + f] = sg0 * fl;
    f1[v * NFQ + p1] = sg1 * fl;
  }
}

// ---- Neighbor_reconcile (Spatial.hpp:716-759) ------------------------------------------------------------
template <int ND, int RS, bool DEF>
__global__ __launch_bounds__(256) void k_neighbor_reconcile(Dev_mesh m)
{
  constexpr int NFQ = Shape<ND, RS>::NFQ, NV = ND + 2;
  constexpr int cls = DEF ? HX_DEFORMED : HX_CARTESIAN;
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int con = gid / NFQ, f = gid % NFQ;
  if (con >= m.n_con[cls]) return;
  const Dir dir(m.con_dir[cls][con]);
  double* f0 = slot_ptr(m, m.con_slot[cls][2 * con]) + NV * NFQ;
  double* f1 = slot_ptr(m, m.con_slot[cls][2 * con + 1]) + NV * NFQ;
  int p1 = f;
  double sg0 = 1., sg1 = 1.;
  if constexpr (DEF) {
    p1 = face1_index<ND, RS>(dir, f);
    sg0 = 1 - 2 * int(dir.flip_normal(0));
    sg1 = 1 - 2 * int(dir.flip_normal(1));
  }
  for (int v = 0; v < NV; ++v) {
    const double a0 = f0[v * NFQ + f], a1 = f1[v * NFQ + p1];
    double avg = 0;
    avg += .5 * sg0 * a0;
    avg += .5 * sg1 * a1;
    f0[v * NFQ + f] = sg0 * avg - a0;
    f1[v * NFQ + p1] = sg1 * avg - a1;
  }
}

// ---- Max_dt (Spatial.hpp:784-828): per-point scale, wavefront + workgroup min, one atomic per workgroup ----
template <int ND, int RS, bool VISC>
__global__ __launch_bounds__(256) void k_max_dt(Dev_mesh m, const Ops<RS>* __restrict__ gops, Ns_params ns, int cls,
                                                int is_local, double max_cfl_c, double max_cfl_d, double* result)
{
  constexpr int NQ = Shape<ND, RS>::NQ, NV = ND + 2, NVERT = 1 << ND;
  __shared__ double wmin[256 / 64];
  const size_t gid = size_t(blockIdx.x) * 256 + threadIdx.x;
  const size_t e = gid / NQ;
  const int q = int(gid % NQ);
  double inv = DBL_MAX;
  if (e < size_t(m.n_elem[cls])) {
    double vs[NVERT];
    for (int i = 0; i < NVERT; ++i) vs[i] = m.vertex_tss[cls][e * NVERT + i];
    // math::interp (math.hpp:207-218)
    int stride = NVERT;
    for (int d = 0; d < ND; ++d) {
      stride /= 2;
      const double x = gops->node[(q / ipow(RS, ND - 1 - d)) % RS];
      for (int i = 0; i < stride; ++i) vs[i] += x * (vs[i + stride] - vs[i]);
    }
    const double spacing = vs[0];
    double u[NV];
    const double* st = m.state[cls] + e * NV * NQ;
    for (int v = 0; v < NV; ++v) u[v] = st[v * NQ + q];
    double scale = char_speed<ND>(u) / max_cfl_c / spacing;
    if constexpr (VISC) {
      double bulk = 0., lap = 0.;
      if (m.av[cls]) {
        bulk = m.av[cls][e * 2 * NQ + q];
        lap = m.av[cls][e * 2 * NQ + NQ + q];
      }
      scale += diffusivity<ND>(ns, u, bulk, lap) / max_cfl_d / spacing / spacing;
    }
    if (is_local) {
      m.tss[cls][e * NQ + q] = 1. / scale;
    } else {
      m.tss[cls][e * NQ + q] = 1.;
      inv = 1. / scale;
    }
  }
  if (is_local) return;
  // NaN never wins, as with std::min(dt, NaN) in the reference
  for (int off = 32; off > 0; off /= 2) {
    const double o = __shfl_xor(inv, off, 64);
    inv = (o < inv) ? o : inv;
  }
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x / 64] = inv;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = wmin[0];
    for (int w = 1; w < 256 / 64; ++w) b = (wmin[w] < b) ? wmin[w] : b;
    if (b < DBL_MAX) atomicMin(reinterpret_cast<unsigned long long*>(result), __double_as_longlong(b));
  }
}

// ---- Prolong_refined / Restrict_refined (Spatial.hpp:153-285) --------------------------------------------
// In-place product of `mat` with every row of a face block along tangential dimension `i_dim`
// (the row loops of Spatial.hpp:183-199 and 259-275).
template <int ND, int RS>
__device__ inline void tangential_matvec(double* var_face, const double (&mat)[RS][RS], int i_dim)
{
  constexpr int NFQ = ipow(RS, ND - 1);
  const int qpoint_stride = ipow(RS, ND - 2 - i_dim);
  for (int i_outer = 0; i_outer < NFQ / (RS * qpoint_stride); ++i_outer)
    for (int i_inner = 0; i_inner < qpoint_stride; ++i_inner) {
      double row[RS], out[RS];
      for (int k = 0; k < RS; ++k) row[k] = var_face[(i_outer * RS + k) * qpoint_stride + i_inner];
      for (int i = 0; i < RS; ++i) {
        double acc = 0;
        for (int k = 0; k < RS; ++k) acc += mat[i][k] * row[k];
        out[i] = acc;
      }
      for (int k = 0; k < RS; ++k) var_face[(i_outer * RS + k) * qpoint_stride + i_inner] = out[k];
    }
}

template <int ND, int RS>
__global__ __launch_bounds__(256) void k_prolong(Dev_mesh m, const Ops<RS>* __restrict__ ops, int n_var, int scale,
                                                 int offset)
{
  constexpr int NFQ = ipow(RS, ND - 1);
  constexpr int N_FACE = 1 << (ND - 1);
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int per = N_FACE * n_var;
  const int i_ref = gid / per;
  if (i_ref >= m.n_ref) return;
  const int i_face = (gid % per) / n_var, v = gid % n_var;
  const bool str[2] = {bool(m.ref_stretch[2 * i_ref]), bool(m.ref_stretch[2 * i_ref + 1])};
  int nf = N_FACE;
  for (int d = 0; d < ND - 1; ++d) nf /= 1 + str[d];
  if (i_face >= nf) return;
  const double* coarse = slot_ptr(m, m.ref_slot[5 * i_ref]) + offset * (ND + 2) * NFQ;
  double* var_face = slot_ptr(m, m.ref_slot[5 * i_ref + 1 + i_face]) + offset * (ND + 2) * NFQ + v * NFQ;
  double buf[NFQ];
  for (int qq = 0; qq < NFQ; ++qq) buf[qq] = coarse[v * NFQ + qq];
  for (int d = 0; d < ND - 1; ++d) {
    if (str[d]) {
      for (int qq = 0; qq < NFQ; ++qq) buf[qq] *= 1 + scale;
    } else {
      const int pw = ND - 2 - d;
      const int face_stride = str[ND - 2] ? 1 : ipow(2, pw);
      const int i_half = (i_face / face_stride) % 2;
      tangential_matvec<ND, RS>(buf, ops->prolong[i_half], d);
    }
  }
  for (int qq = 0; qq < NFQ; ++qq) var_face[qq] = buf[qq];
}

template <int ND, int RS>
__global__ __launch_bounds__(256) void k_restrict(Dev_mesh m, const Ops<RS>* __restrict__ ops, int n_var, int scale,
                                                  int offset)
{
  constexpr int NFQ = ipow(RS, ND - 1);
  constexpr int N_FACE = 1 << (ND - 1);
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int i_ref = gid / n_var, v = gid % n_var;
  if (i_ref >= m.n_ref) return;
  const bool str[2] = {bool(m.ref_stretch[2 * i_ref]), bool(m.ref_stretch[2 * i_ref + 1])};
  int nf = N_FACE;
  for (int d = 0; d < ND - 1; ++d) nf /= 1 + str[d];
  double* coarse = slot_ptr(m, m.ref_slot[5 * i_ref]) + offset * (ND + 2) * NFQ + v * NFQ;
  double acc[NFQ];
  for (int qq = 0; qq < NFQ; ++qq) acc[qq] = 0.;
  for (int i_face = 0; i_face < nf; ++i_face) {
    double* var_face = slot_ptr(m, m.ref_slot[5 * i_ref + 1 + i_face]) + offset * (ND + 2) * NFQ + v * NFQ;
    double buf[NFQ];
    for (int qq = 0; qq < NFQ; ++qq) buf[qq] = var_face[qq];
    for (int d = 0; d < ND - 1; ++d) {
      if (str[d]) {
        for (int qq = 0; qq < NFQ; ++qq) buf[qq] /= 1 + scale;
      } else {
        const int pw = ND - 2 - d;
        const int face_stride = str[ND - 2] ? 1 : ipow(2, pw);
        const int i_half = (i_face / face_stride) % 2;
        tangential_matvec<ND, RS>(buf, ops->restrict_[i_half], d);
      }
    }
    for (int qq = 0; qq < NFQ; ++qq) {
      var_face[qq] = buf[qq];
      acc[qq] += buf[qq];
    }
  }
  for (int qq = 0; qq < NFQ; ++qq) coarse[qq] = acc[qq];
}

// ---- flow boundary conditions (Boundary_condition.cpp) ---------------------------------------------------
template <int ND, int RS>
__device__ inline void reflect_normal(double* g, const double* nrml, int f)
{
  constexpr int NFQ = ipow(RS, ND - 1);
  double dot = 0., norm_sq = 0.;
  for (int d = 0; d < ND; ++d) {
    const double nn = nrml[d * NFQ + f];
    dot += g[d * NFQ + f] * nn;
    norm_sq += nn * nn;
  }
  for (int d = 0; d < ND; ++d) g[d * NFQ + f] -= 2 * dot * nrml[d * NFQ + f] / norm_sq;
}

template <int ND, int RS>
__global__ __launch_bounds__(256) void k_state_bcs(Dev_mesh m)
{
  constexpr int NFQ = ipow(RS, ND - 1), NV = ND + 2;
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int ib = gid / NFQ, f = gid % NFQ;
  if (ib >= m.n_bc_con) return;
  const int con = m.bc_con[ib];
  const hx_bc bc = m.bcs[m.bc_serial[ib]];
  const double* in_f = slot_ptr(m, m.con_slot[HX_DEFORMED][2 * con]);
  double* gh_f = slot_ptr(m, m.con_slot[HX_DEFORMED][2 * con + 1]);
  switch (bc.kind) {
    case HX_BC_FREESTREAM:
      for (int v = 0; v < NV; ++v) gh_f[v * NFQ + f] = bc.state[v];
      break;
    case HX_BC_COPY:
    case HX_BC_OUTFLOW:
      for (int l = 0; l < 2; ++l)
        for (int v = 0; v < NV; ++v) gh_f[(l * NV + v) * NFQ + f] = in_f[(l * NV + v) * NFQ + f];
      break;
    case HX_BC_NONPENETRATION:
      for (int v = 0; v < NV; ++v) gh_f[v * NFQ + f] = in_f[v * NFQ + f];
      reflect_normal<ND, RS>(gh_f, m.normals + size_t(m.con_nslot[2 * con]) * ND * NFQ, f);
      break;
    default:
      break;
  }
}

template <int ND, int RS>
__global__ __launch_bounds__(256) void k_flux_bcs(Dev_mesh m)
{
  constexpr int NFQ = ipow(RS, ND - 1), NV = ND + 2;
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int ib = gid / NFQ, f = gid % NFQ;
  if (ib >= m.n_bc_con) return;
  const int con = m.bc_con[ib];
  const hx_bc bc = m.bcs[m.bc_serial[ib]];
  const double* in_f = slot_ptr(m, m.con_slot[HX_DEFORMED][2 * con]);
  double* gh_f = slot_ptr(m, m.con_slot[HX_DEFORMED][2 * con + 1]);
  switch (bc.kind) {
    case HX_BC_FREESTREAM:
    case HX_BC_COPY:
      for (int l = 0; l < 2; ++l)
        for (int v = 0; v < NV; ++v) gh_f[(l * NV + v) * NFQ + f] = in_f[(l * NV + v) * NFQ + f];
      break;
    case HX_BC_OUTFLOW:
      for (int v = 0; v < NV; ++v) gh_f[(NV + v) * NFQ + f] = -in_f[(NV + v) * NFQ + f];
      break;
    case HX_BC_NONPENETRATION:
      for (int v = 0; v < NV; ++v) gh_f[(NV + v) * NFQ + f] = -in_f[(NV + v) * NFQ + f];
      reflect_normal<ND, RS>(gh_f + NV * NFQ, m.normals + size_t(m.con_nslot[2 * con]) * ND * NFQ, f);
      break;
    default:
      break;
  }
}

// ---- launchers -------------------------------------------------------------------------------------------
inline unsigned blocks_for(size_t n_items, int per_block) { return unsigned((n_items + per_block - 1) / per_block); }

template <int ND, int RS>
struct Launchers
{
  using S = Shape<ND, RS>;
  static const Ops<RS>* ops(const Launch& l) { return static_cast<const Ops<RS>*>(l.ops); }

  static hipError_t write_face(const Launch& l, int cls)
  {
    const int n = l.m.n_elem[cls];
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL((k_write_face<ND, RS>), dim3(blocks_for(n, S::EPB)), dim3(S::BLOCK), 0, l.stream, l.m, ops(l), cls);
    return hipGetLastError();
  }
  static hipError_t neighbor(const Launch& l, int cls, bool visc)
  {
    const int n = l.m.n_con[cls];
    if (n == 0) return hipSuccess;
    dim3 g(blocks_for(size_t(n) * S::NFQ, 256)), b(256);
    if (cls == HX_DEFORMED) {
      if (visc) hipLaunchKernelGGL((k_neighbor<ND, RS, true, true>), g, b, 0, l.stream, l.m);
      else hipLaunchKernelGGL((k_neighbor<ND, RS, true, false>), g, b, 0, l.stream, l.m);
    } else {
      if (visc) hipLaunchKernelGGL((k_neighbor<ND, RS, false, true>), g, b, 0, l.stream, l.m);
      else hipLaunchKernelGGL((k_neighbor<ND, RS, false, false>), g, b, 0, l.stream, l.m);
    }
    return hipGetLastError();
  }
  static hipError_t local(const Launch& l, int cls, bool visc, double coef, int stage, int cr, int uf)
  {
    const int n = l.m.n_elem[cls];
    if (n == 0) return hipSuccess;
    dim3 g(blocks_for(n, S::EPB)), b(S::BLOCK);
    if (visc) {
      if (cls == HX_DEFORMED) hipLaunchKernelGGL((k_local_ns<ND, RS, true>), g, b, 0, l.stream, l.m, ops(l), l.ns, coef, cr, uf);
      else hipLaunchKernelGGL((k_local_ns<ND, RS, false>), g, b, 0, l.stream, l.m, ops(l), l.ns, coef, cr, uf);
    } else {
      if (cls == HX_DEFORMED) hipLaunchKernelGGL((k_local_euler<ND, RS, true>), g, b, 0, l.stream, l.m, ops(l), coef, stage, cr, uf);
      else hipLaunchKernelGGL((k_local_euler<ND, RS, false>), g, b, 0, l.stream, l.m, ops(l), coef, stage, cr, uf);
    }
    return hipGetLastError();
  }
  static hipError_t neighbor_reconcile(const Launch& l, int cls)
  {
    const int n = l.m.n_con[cls];
    if (n == 0) return hipSuccess;
    dim3 g(blocks_for(size_t(n) * S::NFQ, 256)), b(256);
    if (cls == HX_DEFORMED) hipLaunchKernelGGL((k_neighbor_reconcile<ND, RS, true>), g, b, 0, l.stream, l.m);
    else hipLaunchKernelGGL((k_neighbor_reconcile<ND, RS, false>), g, b, 0, l.stream, l.m);
    return hipGetLastError();
  }
  static hipError_t reconcile_ldg_flux(const Launch& l, int cls, double dt, int cr, int uf)
  {
    const int n = l.m.n_elem[cls];
    if (n == 0) return hipSuccess;
    dim3 g(blocks_for(n, S::EPB)), b(S::BLOCK);
    if (cls == HX_DEFORMED) hipLaunchKernelGGL((k_reconcile_ldg<ND, RS, true>), g, b, 0, l.stream, l.m, ops(l), dt, cr, uf);
    else hipLaunchKernelGGL((k_reconcile_ldg<ND, RS, false>), g, b, 0, l.stream, l.m, ops(l), dt, cr, uf);
    return hipGetLastError();
  }
  static hipError_t max_dt(const Launch& l, int cls, bool visc, int is_local, double cc, double cd)
  {
    const int n = l.m.n_elem[cls];
    if (n == 0) return hipSuccess;
    dim3 g(blocks_for(size_t(n) * S::NQ, 256)), b(256);
    if (visc) hipLaunchKernelGGL((k_max_dt<ND, RS, true>), g, b, 0, l.stream, l.m, ops(l), l.ns, cls, is_local, cc, cd, l.scalar);
    else hipLaunchKernelGGL((k_max_dt<ND, RS, false>), g, b, 0, l.stream, l.m, ops(l), l.ns, cls, is_local, cc, cd, l.scalar);
    return hipGetLastError();
  }
  static hipError_t prolong(const Launch& l, int n_var, int scale, int offset)
  {
    if (l.m.n_ref == 0) return hipSuccess;
    hipLaunchKernelGGL((k_prolong<ND, RS>), dim3(blocks_for(size_t(l.m.n_ref) * (1 << (ND - 1)) * n_var, 256)), dim3(256), 0,
                       l.stream, l.m, ops(l), n_var, scale, offset);
    return hipGetLastError();
  }
  static hipError_t restrict_(const Launch& l, int n_var, int scale, int offset)
  {
    if (l.m.n_ref == 0) return hipSuccess;
    hipLaunchKernelGGL((k_restrict<ND, RS>), dim3(blocks_for(size_t(l.m.n_ref) * n_var, 256)), dim3(256), 0, l.stream, l.m,
                       ops(l), n_var, scale, offset);
    return hipGetLastError();
  }
  static hipError_t state_bcs(const Launch& l)
  {
    if (l.m.n_bc_con == 0) return hipSuccess;
    hipLaunchKernelGGL((k_state_bcs<ND, RS>), dim3(blocks_for(size_t(l.m.n_bc_con) * S::NFQ, 256)), dim3(256), 0, l.stream, l.m);
    return hipGetLastError();
  }
  static hipError_t flux_bcs(const Launch& l)
  {
    if (l.m.n_bc_con == 0) return hipSuccess;
    hipLaunchKernelGGL((k_flux_bcs<ND, RS>), dim3(blocks_for(size_t(l.m.n_bc_con) * S::NFQ, 256)), dim3(256), 0, l.stream, l.m);
    return hipGetLastError();
  }
  static const Kernel_set* set()
  {
    static const Kernel_set s{&write_face, &neighbor, &local, &neighbor_reconcile, &reconcile_ldg_flux,
                              &max_dt, &prolong, &restrict_, &state_bcs, &flux_bcs};
    return &s;
  }
};

}  // namespace hx

#define HX_DEFINE_KERNEL_SET(ND)                                   \
  namespace hx {                                                   \
  const Kernel_set* kernel_set_nd##ND(int rs)                      \
  {                                                                \
    switch (rs) {                                                  \
      case 2: return Launchers<ND, 2>::set();                      \
      case 3: return Launchers<ND, 3>::set();                      \
      case 4: return Launchers<ND, 4>::set();                      \
      case 5: return Launchers<ND, 5>::set();                      \
      case 6: return Launchers<ND, 6>::set();                      \
      case 7: return Launchers<ND, 7>::set();                      \
      case 8: return Launchers<ND, 8>::set();                      \
      default: return nullptr;                                     \
    }                                                              \
  }                                                                \
  }
