// gfx950 kernels of the artificial-viscosity generator: pde::Advection (pde.hpp:265-349) and pde::Smooth_art_visc
// (pde.hpp:355-431) through the Spatial kernels, their boundary hooks (Boundary_condition.cpp:12-60,346-367) and the
// point loops of Solver::update_art_visc_smoothness (Solver.cpp:457-582). Included by hx_kernels.hpp.
//
// Element data: the Advection state is the momentum slots of `state` (holding velocity while the generator runs,
// Solver.cpp:465-477) followed by the rs `advection` arrays; Smooth_art_visc works on the 4 `forcing` arrays
// ([0] the derivative estimate, [1..3] the diffused chain). Face slots carry the nd + rs Advection variables (or the
// 3 forcing variables) in block 0 and the LDG variables at offset nv*nfq, as for the flow.
#pragma once

namespace hx {

constexpr int av_kind_adv = HX_PDE_ADVECTION;
constexpr int av_kind_sav = HX_PDE_SMOOTH_AV;
constexpr int av_kind_fta = HX_PDE_FIX_ADMIS;  // pde::Fix_therm_admis: diffusion of the flow state (pde.hpp:437-493)

// Point constants of the generator, passed by value
template <int RS>
struct Av_consts
{
  double node[RS];     // Gauss-Legendre nodes on [0, 1]
  double adv_node[RS]; // Advection's 2 x - 1 (pde.hpp:279-281)
  double weight[RS];   // quadrature weights
  double orth[RS];     // normalized Legendre polynomial of degree av_rs - 1 at the nodes (Solver.cpp:529)
};

template <int ND, int RS, int KIND>
struct Av_vars
{
  static constexpr int N = KIND == av_kind_adv ? ND + RS : KIND == av_kind_sav ? 3 : ND + 2;  // n_extrap
};

// The variables a diffusion-only PDE works on, [N][nq] contiguous per element: forcing[1..3] for Smooth_art_visc
// (pde.hpp:373-378), the conserved state for Fix_therm_admis (pde.hpp:448-453)
template <int ND, int RS, int KIND>
__device__ inline double* diff_vars(const Dev_mesh& m, int cls, size_t e)
{
  constexpr int NQ = Shape<ND, RS>::NQ;
  if constexpr (KIND == av_kind_sav) return m.forcing[cls] + (e * 4 + 1) * NQ;
  else return m.state[cls] + e * (ND + 2) * NQ;
}

// Advection::fetch_extrap (pde.hpp:284-290) / Smooth_art_visc::fetch_extrap (pde.hpp:373-378) of variable v
template <int ND, int RS, int KIND>
__device__ inline double av_extrap(const Dev_mesh& m, int cls, size_t e, int v, int q)
{
  constexpr int NQ = Shape<ND, RS>::NQ;
  if constexpr (KIND == av_kind_adv) {
    return v < ND ? m.state[cls][(e * (ND + 2) + v) * NQ + q] : m.advection[cls][(e * RS + (v - ND)) * NQ + q];
  } else {
    return diff_vars<ND, RS, KIND>(m, cls, e)[v * NQ + q];
  }
}

// ---- Write_face<Advection / Smooth_art_visc> (Spatial.hpp:41-57) ------------------------------------------
template <int ND, int RS, int KIND>
__global__ __launch_bounds__((Shape<ND, RS>::BLOCK)) void k_av_write_face(Dev_mesh m, Eops<RS, double> o, int cls)
{
  using S = Shape<ND, RS>;
  constexpr int NQ = S::NQ, NQP = S::NQP, EPB = S::EPB, NVAR = Av_vars<ND, RS, KIND>::N;
  __shared__ double lds[EPB * NVAR * NQP];
  const int t = threadIdx.x, el = t / NQ, q = t % NQ;
  const int e = blockIdx.x * EPB + el;
  const bool valid = e < m.n_elem[cls];
  for (int v = 0; v < NVAR; ++v)
    lds[el * NVAR * NQP + v * NQP + S::pad(q)] = valid ? av_extrap<ND, RS, KIND>(m, cls, e, v, q) : 0.;
  __syncthreads();
  write_faces_from<ND, RS, S::BLOCK, NVAR>(m, cls, blockIdx.x, o, lds, NVAR * NQP);
}

// ---- Neighbor<Advection / Smooth_art_visc / Fix_therm_admis> (Spatial.hpp:613-704) -----------------------------
// Advection: LLF flux of the rs advected scalars with flux node_i (v.n) a_i and characteristic speed max(1, |v|)
// (pde.hpp:319-347), written over the first rs variables of block 0. Smooth_art_visc and Fix_therm_admis: diffusion
// only, the LDG average of the extrapolated variables.
template <int ND, int RS, bool DEF, int KIND>
__global__ __launch_bounds__(256) void k_av_neighbor(Dev_mesh m, Av_consts<RS> ac)
{
  constexpr int NFQ = Shape<ND, RS>::NFQ, NV = ND + 2, NVAR = Av_vars<ND, RS, KIND>::N;
  constexpr int cls = DEF ? HX_DEFORMED : HX_CARTESIAN;
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int con = gid / NFQ, f = gid % NFQ;
  if (con >= m.n_con[cls]) return;
  const Dir dir(m.con_dir[cls][con]);
  double* f0 = slot_ptr(m, m.con_slot[cls][2 * con]);
  double* f1 = slot_ptr(m, m.con_slot[cls][2 * con + 1]);
  const int p1 = DEF ? face1_index<ND, RS>(dir, f) : f;
  if constexpr (KIND != av_kind_adv) {
    for (int v = 0; v < NVAR; ++v) {
      const double avg = .5 * (f0[v * NFQ + f] + f1[v * NFQ + p1]);
      f0[NV * NFQ + v * NFQ + f] = avg;
      f1[NV * NFQ + v * NFQ + p1] = avg;
    }
  } else {
    double sg0 = 1., sg1 = 1.;
    double n[ND];
    if constexpr (DEF) {
      sg0 = 1 - 2 * int(dir.flip_normal(0));
      sg1 = 1 - 2 * int(dir.flip_normal(1));
      const double* nn = m.normals + size_t(m.con_nslot[2 * con]) * ND * NFQ;
      for (int j = 0; j < ND; ++j) n[j] = sg0 * nn[j * NFQ + f];
    } else {
      for (int j = 0; j < ND; ++j) n[j] = (j == dir.i_dim0);
    }
    double vn0 = 0., vn1 = 0., vsq0 = 0., vsq1 = 0.;
    for (int j = 0; j < ND; ++j) {
      const double v0 = f0[j * NFQ + f], v1 = f1[j * NFQ + p1];
      vn0 += v0 * n[j];
      vn1 += v1 * n[j];
      vsq0 += v0 * v0;
      vsq1 += v1 * v1;
    }
    double a0[RS], a1[RS];
    for (int i = 0; i < RS; ++i) {
      a0[i] = f0[(ND + i) * NFQ + f];
      a1[i] = f1[(ND + i) * NFQ + p1];
    }
    double nsq = 0;
    for (int j = 0; j < ND; ++j) nsq += n[j] * n[j];
    const double c = fmax(fmax(1., sqrt(vsq0)), fmax(1., sqrt(vsq1))) * sqrt(nsq);
    for (int i = 0; i < RS; ++i) {
      const double fl = .5 * (ac.adv_node[i] * vn0 * a0[i] + ac.adv_node[i] * vn1 * a1[i] + c * (a0[i] - a1[i]));
      f0[i * NFQ + f] = sg0 * fl;
      f1[i * NFQ + p1] = sg1 * fl;
    }
  }
}

// ---- Local<Advection> (Spatial.hpp:326-509) -----------------------------------------------------------------
// Convective rate along each direction, the constant source 2/advect_length in stage 0, and Advection::write_update
// (pde.hpp:292-300): stage 0 a = (a + du)/(1 + tss 2/L), stage 1 a += du/(1 + tss 2/L). Then Write_face.
template <int ND, int RS>
struct Adv_lds
{
  using S = Shape<ND, RS>;
  static constexpr int F = 0;                                  // [RS][NQP] flux along the current direction
  static constexpr int RT = F + RS * S::NQP;                   // [RS][NQP] rate
  static constexpr int U = 0;                                  // [ND + RS][NQP] updated state (after F, RT)
  static constexpr int SB = (2 * RS > ND + RS ? 2 * RS : ND + RS) * S::NQP;  // [2][RS][NFQ] the direction's face fluxes
  static constexpr int PER = SB + 2 * RS * S::NFQ;
};

template <int ND, int RS, bool DEF>
__global__ __launch_bounds__((Shape<ND, RS>::BLOCK)) void k_adv_local(Dev_mesh m, Eops<RS, double> o, Av_consts<RS> ac,
                                                                      double coef, int stage, int compute_residual,
                                                                      double advect_length, int project)
{
  using S = Shape<ND, RS>;
  using L = Adv_lds<ND, RS>;
  constexpr int NQ = S::NQ, NQP = S::NQP, NFQ = S::NFQ, EPB = S::EPB, NT = S::BLOCK;
  constexpr int cls = DEF ? HX_DEFORMED : HX_CARTESIAN;
  __shared__ double lds[EPB * L::PER];
  const int t = threadIdx.x, el = t / NQ, q = t % NQ, qp = S::pad(q);
  const int n = m.n_elem[cls];
  const int e = blockIdx.x * EPB + el;
  const bool valid = e < n;
  double* my = lds + el * L::PER;
  const Contraction<RS, Fp64_form> con(o, nullptr);

  double vel[ND], a[RS];
  for (int j = 0; j < ND; ++j) vel[j] = valid ? m.state[cls][(size_t(e) * (ND + 2) + j) * NQ + q] : 0.;
  for (int i = 0; i < RS; ++i) a[i] = valid ? m.advection[cls][(size_t(e) * RS + i) * NQ + q] : 0.;
  double det = 1.;
  double met[ND][ND];
  if constexpr (DEF) {
    const double* mt = m.metrics + size_t(e) * (ND * ND + 1) * NQ;
    for (int i = 0; i < ND; ++i)
      for (int j = 0; j < ND; ++j) met[i][j] = valid ? mt[(i * ND + j) * NQ + q] : double(i == j);
    det = valid ? mt[ND * ND * NQ + q] : 1.;
  }
  // numerical face fluxes (the first rs variables of block 0 of every face), loaded now into registers, direction by
  // direction, and staged into LDS one direction at a time: entry k of direction i's block [EPB][2][RS][NFQ] is held by
  // thread k mod NT
  constexpr int DB = EPB * 2 * RS * NFQ, FPD = (DB + NT - 1) / NT;
  double sbr[ND][FPD];
#pragma unroll
  for (int i = 0; i < ND; ++i)
#pragma unroll
    for (int j = 0; j < FPD; ++j) {
      const int k = t + j * NT, el2 = k / (2 * RS * NFQ), r = k % (2 * RS * NFQ);
      const int e2 = blockIdx.x * EPB + el2;
      sbr[i][j] = k < DB && e2 < n ? slot_ptr(m, m.face_slot[cls][e2 * 2 * ND + 2 * i + r / (RS * NFQ)])[r % (RS * NFQ)]
                                   : 0.;
    }
  for (int i = 0; i < ND; ++i) {
    double vn = 0;
    for (int j = 0; j < ND; ++j) {
      if constexpr (DEF) vn += vel[j] * met[i][j];
      else vn += vel[j] * double(j == i);
    }
    if (i) __syncthreads();
    for (int v = 0; v < RS; ++v) my[L::F + v * NQP + qp] = ac.adv_node[v] * vn * a[v];
#pragma unroll
    for (int j = 0; j < FPD; ++j) {
      const int k = t + j * NT;
      if (k < DB) lds[(k / (2 * RS * NFQ)) * L::PER + L::SB + k % (2 * RS * NFQ)] = sbr[i][j];
    }
    __syncthreads();
    for_columns<ND, RS, RS, NT, false>(i, [&](bool, int el2, int v, int r, int base, int sd) {
      double* b = lds + el2 * L::PER;
      double x[RS], y[RS];
      load_row(b + L::F + v * NQP + base, sd, x);
      con.faces(x, b[L::SB + v * NFQ + r], b[L::SB + (RS + v) * NFQ + r], y);
      sub_row(b + L::RT + v * NQP + base, sd, y, i == 0);
    });
  }
  __syncthreads();
  if (valid) {
    const double d_pos = m.nom_size[cls][e];
    const double tss = m.tss[cls][size_t(e) * NQ + q];
    double mult = coef * tss / d_pos;
    if constexpr (DEF) mult /= det;
    double src_mult = d_pos;
    if constexpr (DEF) src_mult *= det;
    double* rc = m.rcache[cls] + size_t(e) * S::NRC * NQ;
    const double pseudo = 1 + tss * 2 / advect_length;
    for (int v = 0; v < RS; ++v) {
      double u = my[L::RT + v * NQP + qp];
      if (stage) {
        u -= rc[v * NQ + q];
      } else {
        rc[v * NQ + q] = u;
        u += src_mult * (2 / advect_length);
      }
      u *= mult;
      double du = 0.;
      if (compute_residual) rc[v * NQ + q] = u;
      else du = u;
      if (!stage) a[v] = (a[v] + du) / pseudo;
      else a[v] += du / pseudo;
      m.advection[cls][(size_t(e) * RS + v) * NQ + q] = a[v];
    }
    // project: the advection's last stage also forms the forcing from the advected scalars just written, as the
    // Legendre projection pass after the advection does (Solver.cpp:525-540; ac.orth is of the projection's degree)
    if (project && !compute_residual) {
      double proj = 0;
      for (int i = 0; i < RS; ++i) proj += a[i] * ac.weight[i] * ac.orth[i];
      const double* st = m.state[cls] + size_t(e) * (ND + 2) * NQ;
      m.forcing[cls][size_t(e) * 4 * NQ + q] = proj * proj * 2 * st[(ND + 1) * NQ + q] / st[ND * NQ + q];
    }
  }
  // Write_face<Advection> of the updated state (Spatial.hpp:507)
  __syncthreads();
  for (int j = 0; j < ND; ++j) my[L::U + j * NQP + qp] = vel[j];
  for (int v = 0; v < RS; ++v) my[L::U + (ND + v) * NQP + qp] = a[v];
  __syncthreads();
  write_faces_from<ND, RS, NT, ND + RS>(m, cls, blockIdx.x, o, lds + L::U, L::PER);
}

// ---- Local<Smooth_art_visc / Fix_therm_admis> (Spatial.hpp:326-509, stage 0) -----------------------------------
// LDG gradient of the diffused variables, flux -gradient*normal (pde.hpp:405-410, 479-483), for Smooth_art_visc the
// chained source |f0|, sqrt|f1|, |f2| over diff_time (pde.hpp:419-427), and the additive update of write_update
// (pde.hpp:380-388, 455-458).
template <int ND, int RS, bool DEF, int NS>
struct Sav_lds
{
  using S = Shape<ND, RS>;
  static constexpr int NQP = S::NQP, NFQ = S::NFQ;
  static constexpr int U = 0;                                   // [NS][NQP] the diffused variables
  static constexpr int G = U + NS * NQP;                        // [NS][NQP] one gradient component
  static constexpr int SN = G + NS * NQP;                       // [ND][NQP] reference-level normals n_ij of one j
  static constexpr int SL = SN + (DEF ? ND * NQP : 0);          // [2ND][NS][NFQ] LDG face state
  static constexpr int SFN = SL + 2 * ND * NS * NFQ;            // [2ND][ND][NFQ] face normals
  static constexpr int P1 = SFN + (DEF ? 2 * ND * ND * NFQ : 0);
  static constexpr int F = 0;                                   // [NS][NQP] viscous flux along the current direction
  static constexpr int RT = F + NS * NQP;                       // [NS][NQP] rate
  static constexpr int P2 = RT + NS * NQP;
  static constexpr int PER = P1 > P2 ? P1 : P2;
};

template <int ND, int RS, bool DEF, int KIND>
__global__ __launch_bounds__((Shape<ND, RS>::BLOCK)) void k_sav_local(Dev_mesh m, Eops<RS, double> o, double coef,
                                                                      double diff_time)
{
  using S = Shape<ND, RS>;
  constexpr int NS = Av_vars<ND, RS, KIND>::N;
  constexpr bool SAV = KIND == av_kind_sav;
  using L = Sav_lds<ND, RS, DEF, NS>;
  constexpr int NQ = S::NQ, NQP = S::NQP, NFQ = S::NFQ, NV = S::NV, EPB = S::EPB, NT = S::BLOCK;
  constexpr int cls = DEF ? HX_DEFORMED : HX_CARTESIAN;
  __shared__ double lds[EPB * L::PER];
  const int t = threadIdx.x, el = t / NQ, q = t % NQ, qp = S::pad(q);
  const int n = m.n_elem[cls];
  const int e = blockIdx.x * EPB + el;
  const bool valid = e < n;
  double* my = lds + el * L::PER;
  const Contraction<RS, Fp64_form> con(o, nullptr);

  double uv[NS];
  for (int v = 0; v < NS; ++v) uv[v] = valid ? diff_vars<ND, RS, KIND>(m, cls, e)[v * NQ + q] : 0.;
  for (int v = 0; v < NS; ++v) my[L::U + v * NQP + qp] = uv[v];
  // read with the first loads rather than where they are used: the source's forcing[0] (its forcing[1] and [2] are
  // the first two diffused variables, held above), the update's time step scale, and the elements' face slots (into
  // LDS, for the flux extrapolations)
  const double f0 = (SAV && valid) ? m.forcing[cls][size_t(e) * 4 * NQ + q] : 0.;
  const double tss_q = valid ? m.tss[cls][size_t(e) * NQ + q] : 0.;
  // (the slots are staged where their LDS leaves the workgroups per compute unit as they are)
  constexpr int LDS = 8 * EPB * L::PER, SLOTS = 4 * EPB * 2 * ND;
  constexpr bool STAGE = 163840 / (LDS + SLOTS) == 163840 / LDS;
  __shared__ int fsl[STAGE ? EPB * 2 * ND : 1];
  if (STAGE && t < EPB * 2 * ND) {
    const int e2 = blockIdx.x * EPB + t / (2 * ND);
    fsl[t] = e2 < n ? m.face_slot[cls][e2 * 2 * ND + t % (2 * ND)] : 0;
  }
  double det = 1.;
  double met[ND][ND];
  if constexpr (DEF) {
    const double* mt = m.metrics + size_t(e) * (ND * ND + 1) * NQ;
    for (int i = 0; i < ND; ++i)
      for (int j = 0; j < ND; ++j) met[i][j] = valid ? mt[(i * ND + j) * NQ + q] : double(i == j);
    // the column of the normals that gradient component 0 needs (each later component's replaces it)
    for (int i = 0; i < ND; ++i) my[L::SN + i * NQP + qp] = met[i][0];
    det = valid ? mt[ND * ND * NQ + q] : 1.;
  }
  for (int k = t; k < EPB * 2 * ND * NS * NFQ; k += NT) {
    const int el2 = k / (2 * ND * NS * NFQ), r = k % (2 * ND * NS * NFQ), f = r / (NS * NFQ), vq = r % (NS * NFQ);
    const int e2 = blockIdx.x * EPB + el2;
    lds[el2 * L::PER + L::SL + r] = e2 < n ? slot_ptr(m, m.face_slot[cls][e2 * 2 * ND + f])[NV * NFQ + vq] : 0.;
  }
  if constexpr (DEF) {
    for (int k = t; k < EPB * 2 * ND * ND * NFQ; k += NT) {
      const int el2 = k / (2 * ND * ND * NFQ), r = k % (2 * ND * ND * NFQ), f = r / (ND * NFQ), jq = r % (ND * NFQ);
      const int e2 = blockIdx.x * EPB + el2;
      double val = 0.;
      if (e2 < n) {
        const int ns_ = m.face_nslot[e2 * 2 * ND + f];
        val = ns_ >= 0 ? m.normals[size_t(ns_) * ND * NFQ + jq] : double(jq / NFQ == f / 2);
      }
      lds[el2 * L::PER + L::SFN + r] = val;
    }
  }
  __syncthreads();
  const double d_pos = valid ? m.nom_size[cls][e] : 1.;
  const double inv_h = 1. / d_pos;
  double grad[NS][ND];
  for (int j = 0; j < ND; ++j) {
    if constexpr (DEF) {
      if (j) {
        // (every read of the previous column ended at the barrier after its gradient component)
        for (int i = 0; i < ND; ++i) my[L::SN + i * NQP + qp] = met[i][j];
        __syncthreads();
      }
    }
    for (int i = DEF ? 0 : j; i < (DEF ? ND : j + 1); ++i) {
      for_columns<ND, RS, NS, NT, false>(i, [&](bool, int el2, int v, int r, int base, int sd) {
        double* b = lds + el2 * L::PER;
        double x[RS], y[RS];
        load_row(b + L::U + v * NQP + base, sd, x);
        double b0 = b[L::SL + ((2 * i) * NS + v) * NFQ + r], b1 = b[L::SL + ((2 * i + 1) * NS + v) * NFQ + r];
        if constexpr (DEF) {
          double nk[RS];
          load_row(b + L::SN + i * NQP + base, sd, nk);
#pragma unroll
          for (int k = 0; k < RS; ++k) x[k] *= nk[k];
          b0 *= b[L::SFN + ((2 * i) * ND + j) * NFQ + r];
          b1 *= b[L::SFN + ((2 * i + 1) * ND + j) * NFQ + r];
        }
        con.faces(x, b0, b1, y);
        double* out = b + L::G + v * NQP + base;
#pragma unroll
        for (int k = 0; k < RS; ++k) out[k * sd] = (i == 0 || !DEF) ? y[k] : out[k * sd] + y[k];
      });
      __syncthreads();
    }
    // the column threads above serve several elements; each point scales by its own element's size
    for (int v = 0; v < NS; ++v) grad[v][j] = my[L::G + v * NQP + qp] * inv_h;
    __syncthreads();
  }
  if constexpr (DEF)
    for (int v = 0; v < NS; ++v)
      for (int j = 0; j < ND; ++j) grad[v][j] /= det;
  if constexpr (SAV) {
    double src_mult = d_pos;
    if constexpr (DEF) src_mult *= det;
    for (int v = 0; v < NS; ++v) {
      const double f = valid ? fabs(v == 0 ? f0 : uv[v - 1]) : 0.;
      my[L::RT + v * NQP + qp] = src_mult * (((v == 1) ? sqrt(f) : f) / diff_time);
    }
  } else {
    for (int v = 0; v < NS; ++v) my[L::RT + v * NQP + qp] = 0.;
  }
  for (int i = 0; i < ND; ++i) {
    if (i) __syncthreads();
    for (int v = 0; v < NS; ++v) {
      double acc = 0;
      for (int j = 0; j < ND; ++j) {
        if constexpr (DEF) acc += -grad[v][j] * met[i][j];
        else acc += -grad[v][j] * double(j == i);
      }
      my[L::F + v * NQP + qp] = acc;
    }
    __syncthreads();
    for_columns<ND, RS, NS, NT, false>(i, [&](bool, int el2, int v, int r, int base, int sd) {
      double* b = lds + el2 * L::PER;
      double x[RS], y[RS], bx[2];
      load_row(b + L::F + v * NQP + base, sd, x);
      con.faces_if(false, x, 0., 0., y, bx);
      const int e2 = blockIdx.x * EPB + el2;
      if (e2 < n) {
        const int* fs = (STAGE ? fsl + el2 * 2 * ND : m.face_slot[cls] + e2 * 2 * ND) + 2 * i;
        slot_ptr(m, fs[0])[(NV + v) * NFQ + r] = bx[0];
        slot_ptr(m, fs[1])[(NV + v) * NFQ + r] = bx[1];
      }
      sub_row(b + L::RT + v * NQP + base, sd, y, false);
    });
  }
  __syncthreads();
  if (valid) {
    double mult = coef * tss_q / d_pos;
    if constexpr (DEF) mult /= det;
    double* out = diff_vars<ND, RS, KIND>(m, cls, e);
    for (int v = 0; v < NS; ++v) {
      double u = 0.;
      u += my[L::RT + v * NQP + qp];
      u *= mult;
      out[v * NQ + q] = uv[v] + u;
    }
  }
}

// ---- Reconcile_ldg_flux<Smooth_art_visc / Fix_therm_admis> (Spatial.hpp:543-594) --------------------------------
// -boundary_term(jump), the critical write_update (Smooth_art_visc: f += du; f /= 1 + tss cheby/diff_time,
// pde.hpp:380-388; Fix_therm_admis: u += du, pde.hpp:455-458), then Write_face of the diffused variables.
template <int ND, int RS, bool DEF, int KIND>
__global__ __launch_bounds__((Shape<ND, RS>::BLOCK)) void k_sav_reconcile(Dev_mesh m, Eops<RS, double> o, double dt,
                                                                          double diff_time, double cheby)
{
  using S = Shape<ND, RS>;
  using R = Row<ND, RS>;
  constexpr int NS = Av_vars<ND, RS, KIND>::N;
  constexpr int NQ = S::NQ, NQP = S::NQP, NFQ = S::NFQ, NV = S::NV, EPB = S::EPB;
  constexpr int cls = DEF ? HX_DEFORMED : HX_CARTESIAN;
  constexpr int PF = 2 * ND * NS * NFQ, PER = PF + NS * NQP;
  __shared__ double lds[EPB * PER];
  const int t = threadIdx.x, el = t / NQ, q = t % NQ;
  const int n = m.n_elem[cls];
  const int e = blockIdx.x * EPB + el;
  const bool valid = e < n;
  double* my = lds + el * PER;
  // the point's own data are read with the face data, so that their latency overlaps it
  double fv[NS];
  for (int v = 0; v < NS; ++v) fv[v] = valid ? diff_vars<ND, RS, KIND>(m, cls, e)[v * NQ + q] : 0.;
  const double tss = valid ? m.tss[cls][size_t(e) * NQ + q] : 0., nom = valid ? m.nom_size[cls][e] : 1.;
  const double det = (DEF && valid) ? m.metrics[size_t(e) * (ND * ND + 1) * NQ + ND * ND * NQ + q] : 1.;
  for (int k = t; k < EPB * PF; k += S::BLOCK) {
    const int el2 = k / PF, r = k % PF, f = r / (NS * NFQ), vq = r % (NS * NFQ);
    const int e2 = blockIdx.x * EPB + el2;
    lds[el2 * PER + r] = e2 < n ? slot_ptr(m, m.face_slot[cls][e2 * 2 * ND + f])[NV * NFQ + vq] : 0.;
  }
  __syncthreads();
  double rate[NS] = {};
  for (int i = 0; i < ND; ++i) {
    const int ki = R::node(q, i), fq = R::fq(q, i);
    for (int v = 0; v < NS; ++v)
      rate[v] -= o.lift[ki][0] * my[((2 * i) * NS + v) * NFQ + fq] + o.lift[ki][1] * my[((2 * i + 1) * NS + v) * NFQ + fq];
  }
  if (valid) {
    double mult = dt * tss / nom;
    if constexpr (DEF) mult /= det;
    double* out = diff_vars<ND, RS, KIND>(m, cls, e);
    for (int v = 0; v < NS; ++v) {
      fv[v] += rate[v] * mult;
      if constexpr (KIND == av_kind_sav) fv[v] /= 1 + tss * cheby / diff_time;
      out[v * NQ + q] = fv[v];
    }
  }
  for (int v = 0; v < NS; ++v) my[PF + v * NQP + S::pad(q)] = fv[v];
  __syncthreads();
  write_faces_from<ND, RS, S::BLOCK, NS>(m, cls, blockIdx.x, o, lds + PF, PER);
}

// ---- Max_dt<Advection / Smooth_art_visc> (Spatial.hpp:784-828) ------------------------------------------------
template <int ND, int RS, bool SMOOTH>
__global__ __launch_bounds__(256) void k_av_max_dt(Dev_mesh m, Av_consts<RS> ac, int cls, int is_local,
                                                   double max_cfl_c, double max_cfl_d, double* result)
{
  constexpr int NQ = Shape<ND, RS>::NQ, NVERT = 1 << ND;
  __shared__ double wmin[256 / 64];
  const size_t n_items = size_t(m.n_elem[cls]) * NQ;
  double inv = DBL_MAX;
  for (size_t gid = size_t(blockIdx.x) * 256 + threadIdx.x; gid < n_items; gid += size_t(gridDim.x) * 256) {
    const size_t e = gid / NQ;
    const int q = int(gid % NQ);
    double vs[NVERT];
    for (int i = 0; i < NVERT; ++i) vs[i] = m.vertex_tss[cls][e * NVERT + i];
    int stride = NVERT;
    for (int d = 0; d < ND; ++d) {
      stride /= 2;
      const double x = ac.node[(q / ipow(RS, ND - 1 - d)) % RS];
      for (int i = 0; i < stride; ++i) vs[i] += x * (vs[i + stride] - vs[i]);
    }
    const double spacing = vs[0];
    double scale = 0;
    if constexpr (SMOOTH) {
      scale += 1. / max_cfl_d / spacing / spacing;
    } else {
      double vsq = 0;
      for (int j = 0; j < ND; ++j) {
        const double v = m.state[cls][(e * (ND + 2) + j) * NQ + q];
        vsq += v * v;
      }
      scale += fmax(1., sqrt(vsq)) / max_cfl_c / spacing;
    }
    if (is_local) {
      m.tss[cls][e * NQ + q] = 1. / scale;
    } else {
      m.tss[cls][e * NQ + q] = 1.;
      const double c = 1. / scale;
      inv = (c < inv) ? c : inv;
    }
  }
  if (is_local) return;
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

// ---- boundary hooks (Solver.cpp:83-101,508-514; Boundary_condition.cpp:12-60,346-367) -------------------------
// mode 0: apply_advection; 1: apply_diffusion (ghost state = inside); 2: flux_diffusion (ghost LDG = -inside)
template <int ND, int RS>
__global__ __launch_bounds__(256) void k_av_bcs(Dev_mesh m, int mode)
{
  constexpr int NFQ = Shape<ND, RS>::NFQ, NV = ND + 2;
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int i = gid / NFQ, f = gid % NFQ;
  if (i >= m.n_bc_con) return;
  const int con = m.bc_con[i];
  const double* in_f = slot_ptr(m, m.con_slot[HX_DEFORMED][2 * con]);
  double* gh_f = slot_ptr(m, m.con_slot[HX_DEFORMED][2 * con + 1]);
  if (mode == 1) {
    for (int v = 0; v < NV; ++v) gh_f[v * NFQ + f] = in_f[v * NFQ + f];
    return;
  }
  if (mode == 2) {
    for (int v = 0; v < NV; ++v) gh_f[(NV + v) * NFQ + f] = -in_f[(NV + v) * NFQ + f];
    return;
  }
  const int kind = m.bcs[m.bc_serial[i]].kind;
  if (kind == HX_BC_COPY) {  // copy_state: the first nv variables of both blocks
    for (int l = 0; l < 2; ++l)
      for (int v = 0; v < NV; ++v) gh_f[(l * NV + v) * NFQ + f] = in_f[(l * NV + v) * NFQ + f];
  } else if (kind == HX_BC_NONPENETRATION) {
    const double* nrml = m.normals + size_t(m.con_nslot[2 * con]) * ND * NFQ;
    double dot = 0., nsq = 0.;
    for (int d = 0; d < ND; ++d) {
      dot += in_f[d * NFQ + f] * nrml[d * NFQ + f];
      nsq += nrml[d * NFQ + f] * nrml[d * NFQ + f];
    }
    for (int d = 0; d < ND; ++d) gh_f[d * NFQ + f] = in_f[d * NFQ + f] - 2 * dot * nrml[d * NFQ + f] / nsq;
    for (int a = 0; a < RS; ++a) gh_f[(ND + a) * NFQ + f] = in_f[(ND + a) * NFQ + f];
  } else if (kind == HX_BC_NO_SLIP) {  // Boundary_condition.cpp:429-448: velocity reversed, scalars kept
    for (int d = 0; d < ND; ++d) gh_f[d * NFQ + f] = -in_f[d * NFQ + f];
    for (int a = 0; a < RS; ++a) gh_f[(ND + a) * NFQ + f] = in_f[(ND + a) * NFQ + f];
  } else {  // Flow_bc::apply_advection: velocity copied, advected scalars 2 - inside
    for (int d = 0; d < ND; ++d) gh_f[d * NFQ + f] = in_f[d * NFQ + f];
    for (int a = 0; a < RS; ++a) gh_f[(ND + a) * NFQ + f] = 2. - in_f[(ND + a) * NFQ + f];
  }
}

// ---- Solver::is_admissible (Solver.cpp:921-959) with thermo::admissible (thermo.cpp:6-18) ----------------------
// Density and energy must be positive at every quadrature point of the element and of its faces (state block);
// flags[0] |= 1 when some element is not, flags[1] |= 1 when a checked value is not finite (the reference's "state is
// not finite" assertion). Per element the reference stops at the first inadmissible array: the state, then the faces
// in order; a face after that one is not checked for finiteness.
template <int ND, int RS>
__global__ __launch_bounds__((Shape<ND, RS>::BLOCK)) void k_admissible(Dev_mesh m, int cls, int* flags)
{
  using S = Shape<ND, RS>;
  constexpr int NQ = S::NQ, NFQ = S::NFQ, NV = S::NV, EPB = S::EPB, NF = 2 * ND;
  __shared__ int bad[EPB], nonfinite[EPB], face_bad[EPB], face_nonfinite[EPB];
  const int t = threadIdx.x, el = t / NQ, q = t % NQ;
  const int e = blockIdx.x * EPB + el;
  const bool valid = e < m.n_elem[cls];
  if (q == 0) bad[el] = nonfinite[el] = face_bad[el] = face_nonfinite[el] = 0;
  __syncthreads();
  if (valid) {
    const double* st = m.state[cls] + size_t(e) * NV * NQ;
    bool fin = true;
    for (int v = 0; v < NV; ++v) fin = fin && isfinite(st[v * NQ + q]);
    if (!(st[ND * NQ + q] > 0.) || !(st[(ND + 1) * NQ + q] > 0.)) atomicOr(&bad[el], 1);
    if (!fin) atomicOr(&nonfinite[el], 1);
    for (int k = q; k < NF * NFQ; k += NQ) {
      const int f = k / NFQ, p = k % NFQ;
      const double* fc = slot_ptr(m, m.face_slot[cls][e * NF + f]);
      bool ffin = true;
      for (int v = 0; v < NV; ++v) ffin = ffin && isfinite(fc[v * NFQ + p]);
      if (!(fc[ND * NFQ + p] > 0.) || !(fc[(ND + 1) * NFQ + p] > 0.)) atomicOr(&face_bad[el], 1 << f);
      if (!ffin) atomicOr(&face_nonfinite[el], 1 << f);
    }
  }
  __syncthreads();
  if (valid && q == 0) {
    bool admis = !bad[el], fin = !nonfinite[el];
    for (int f = 0; f < NF && admis; ++f) {
      if ((face_nonfinite[el] >> f) & 1) fin = false;
      if ((face_bad[el] >> f) & 1) admis = false;
    }
    if (!admis) atomicOr(flags, 1);
    if (!fin) atomicOr(flags + 1, 1);
  }
}

// The fine faces of every hanging face (Solver.cpp:946-957). The reference's short-circuit there depends on how
// OpenMP splits the loop; here every fine face is checked for finiteness.
template <int ND, int RS>
__global__ __launch_bounds__(256) void k_admissible_ref(Dev_mesh m, int* flags)
{
  constexpr int NFQ = Shape<ND, RS>::NFQ, NV = ND + 2, NH = 1 << (ND - 1);
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int i = gid / (NH * NFQ), r = gid % (NH * NFQ), f = r / NFQ, p = r % NFQ;
  if (i >= m.n_ref) return;
  int n_fine = NH;
  for (int d = 0; d < ND - 1; ++d)
    if (m.ref_stretch[2 * i + d]) n_fine /= 2;
  if (f >= n_fine) return;
  const double* fc = slot_ptr(m, m.ref_slot[5 * i + 1 + f]);
  bool fin = true;
  for (int v = 0; v < NV; ++v) fin = fin && isfinite(fc[v * NFQ + p]);
  if (!(fc[ND * NFQ + p] > 0.) || !(fc[(ND + 1) * NFQ + p] > 0.)) atomicOr(flags, 1);
  if (!fin) atomicOr(flags + 1, 1);
}

// The artificial-viscosity swap around one iteration of Solver::fix_admissibility (Solver.cpp:1040-1047,
// 1080-1086) with the fix coefficient taken as zero. mode 0: laplacian <- bulk, bulk <- 0; mode 1: bulk <- laplacian,
// laplacian <- 0.
template <int ND, int RS>
__global__ __launch_bounds__(256) void k_fix_av(Dev_mesh m, int cls, int mode)
{
  constexpr int NQ = Shape<ND, RS>::NQ;
  const size_t gid = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (gid >= size_t(m.n_elem[cls]) * NQ) return;
  const size_t e = gid / NQ, q = gid % NQ;
  double* bulk = m.av[cls] + e * 2 * NQ + q;
  double* lap = bulk + NQ;
  if (mode == 0) {
    *lap = *bulk;
    *bulk = 0.;
  } else {
    *bulk = *lap;
    *lap = 0.;
  }
}

// ---- point loops of update_art_visc_smoothness (Solver.cpp:525-540,549-574) ------------------------------------
// what 1: Legendre projection of the advected scalars into forcing[0] (only without advection iterations: otherwise
// the last Advection stage's Local forms it, k_adv_local with project); 2: new coefficient us_max f/(us_max + f),
// squared change weighted into `resid`. (The momentum's division by sqrt(2 rho E) before the advection and its
// restoration in the last loop run in the Write_face that follows each: k_write_face with scale_velocity.)
template <int ND, int RS>
__global__ __launch_bounds__(256) void k_av_points(Dev_mesh m, Av_consts<RS> ac, int cls, int what, double mult,
                                                   double us_max, double* resid)
{
  constexpr int NQ = Shape<ND, RS>::NQ, NV = ND + 2;
  __shared__ double wsum[256 / 64];
  // grid-stride over the points (the grid is capped, av_points_blocks), so that the residual takes one atomic per
  // workgroup of a small grid rather than one per 256 points on a single word
  double part = 0.;
  const size_t n_pts = size_t(m.n_elem[cls]) * NQ;
  for (size_t gid = size_t(blockIdx.x) * 256 + threadIdx.x; gid < n_pts; gid += size_t(gridDim.x) * 256) {
    const size_t e = gid / NQ;
    const int q = int(gid % NQ);
    double* st = m.state[cls] + e * NV * NQ;
    if (what == 1) {
      const double* adv = m.advection[cls] + e * RS * NQ;
      double proj = 0;
      for (int i = 0; i < RS; ++i) proj += adv[i * NQ + q] * ac.weight[i] * ac.orth[i];
      m.forcing[cls][e * 4 * NQ + q] = proj * proj * 2 * st[(ND + 1) * NQ + q] / st[ND * NQ + q];
    } else {
      const double f = mult * m.forcing[cls][(e * 4 + 3) * NQ + q];
      const double new_av = us_max * f / (us_max + f);
      double* av = m.av[cls] + e * 2 * NQ;
      double w = 1;
      for (int d = 0; d < ND; ++d) w *= ac.weight[(q / ipow(RS, ND - 1 - d)) % RS];
      double volume = 1;
      for (int d = 0; d < ND; ++d) volume *= m.nom_size[cls][e];
      const double diff = av[q] - new_av;
      part += (1. * diff * diff) * w * volume;
      av[q] = new_av;
    }
  }
  if (what != 2) return;
  for (int off = 32; off > 0; off /= 2) part += __shfl_xor(part, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x / 64] = part;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = 0;
    for (int w = 0; w < 256 / 64; ++w) b += wsum[w];
    atomicAdd(resid, b);
  }
}

// ---- stabilizing artificial viscosity (stabilizing_art_visc.cpp:8-61, Solver.cpp:636-659) --------------------------
// Per element, the specific volume's share of L2 energy in the top Legendre mode along each dimension,
//   nonsmooth = sum_dims sum_rows (row . orth_{rs-1} w)^2 face_weight / (n_dim ||1/rho||^2),
// mapped to [0, 1] by a sine ramp of half width 0.5 around log10 = -4.25 log10(rs - 1), times (rs - 1) char_speed h.
// One thread per point stages 1/rho (and its weighted square) in LDS, one thread per row forms its squared projection,
// and one wavefront per element sums both and writes the value to every vertex slot of the element (out[e][2^nd],
// ready for sharing by maximum).
template <int ND, int RS>
__global__ __launch_bounds__((Shape<ND, RS>::BLOCK)) void k_stab_indicator(Dev_mesh m, Av_consts<RS> c, int cls,
                                                                        double char_speed, double* __restrict__ out)
{
  using S = Shape<ND, RS>;
  using R = Row<ND, RS>;
  constexpr int NQ = S::NQ, NFQ = S::NFQ, EPB = S::EPB, NR = ND * NFQ;
  __shared__ double iv[EPB][NQ];
  __shared__ double pw[EPB][NQ];  // each point's (1/rho)^2 times its quadrature weight
  __shared__ double rowv[EPB][NR];
  const int t = threadIdx.x, el = t / NQ, q = t % NQ;
  const int n = m.n_elem[cls];
  const int e = blockIdx.x * EPB + el;
  const bool valid = el < EPB && e < n;
  if (el < EPB) {
    const double x = valid ? 1. / m.state[cls][size_t(e) * S::NV * NQ + ND * NQ + q] : 0.;
    double qw = 1.;
    for (int d = 0; d < ND; ++d) qw *= c.weight[(q / ipow(RS, ND - 1 - d)) % RS];
    iv[el][q] = x;
    pw[el][q] = x * x * qw;
  }
  __syncthreads();
  for (int k = t; k < EPB * NR; k += S::BLOCK) {
    const int el2 = k / NR, r = k % NR, i_dim = r / NFQ, f = r % NFQ;
    const int base = R::base_of_fq(f, i_dim), sd = R::stride(i_dim);
    double dot = 0.;
    for (int j = 0; j < RS; ++j) dot += iv[el2][base + j * sd] * (c.orth[j] * c.weight[j]);
    double fw = 1.;
    for (int d = 0; d < ND - 1; ++d) fw *= c.weight[(f / ipow(RS, ND - 2 - d)) % RS];
    rowv[el2][r] = dot * dot * fw;
  }
  __syncthreads();
  // one wavefront per element: the point and row sums lane-strided, then across the lanes (a fixed order, not the
  // reference's sequential one; the sums agree to rounding). Only whole wavefronts take part (a block's last one may be
  // partial, and a shuffle must not read a lane the block does not have); every block has at least two.
  constexpr int NWAVE = S::BLOCK / 64;
  static_assert(NWAVE >= 1, "the block holds a whole wavefront");
  const int lane = t & 63;
  for (int el2 = t / 64; el2 < EPB && t / 64 < NWAVE; el2 += NWAVE) {
    const int e2 = blockIdx.x * EPB + el2;
    double norm_sq = 0., nonsmooth = 0.;
    for (int p = lane; p < NQ; p += 64) norm_sq += pw[el2][p];
    for (int r = lane; r < NR; r += 64) nonsmooth += rowv[el2][r];
    for (int off = 32; off > 0; off /= 2) {
      norm_sq += __shfl_xor(norm_sq, off, 64);
      nonsmooth += __shfl_xor(nonsmooth, off, 64);
    }
    if (lane != 0 || e2 >= n) continue;
    nonsmooth /= norm_sq * ND;
    const double center = -4.25 * log(double(RS - 1)) / log(10.), half_width = .5;
    double ind = log(nonsmooth) / log(10.);
    if (ind <= center - half_width) ind = 0.;
    else if (ind < center + half_width) ind = .5 * (1. + sin(M_PI * (ind - center) / 2. / half_width));
    else ind = 1.;
    const double val = (RS - 1) * char_speed * m.nom_size[cls][e2] * ind;
    for (int v = 0; v < (1 << ND); ++v) out[size_t(e2) * (1 << ND) + v] = val;
  }
}

// Laplacian artificial viscosity at every point from the element's vertex values vert[e][2^nd] (first vertex index
// slowest): Gauss_lobatto(2)'s linear interpolant (1 - x, x) at the nodes along each dimension, applied as
// math::hypercube_matvec does (the last dimension contracted first).
template <int ND, int RS>
__global__ __launch_bounds__(256) void k_vertex_interp_lap(Dev_mesh m, Av_consts<RS> c, int cls,
                                                           const double* __restrict__ vert)
{
  constexpr int NQ = Shape<ND, RS>::NQ, NVERT = 1 << ND;
  const size_t gid = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (gid >= size_t(m.n_elem[cls]) * NQ) return;
  const size_t e = gid / NQ;
  const int q = int(gid % NQ);
  double v[NVERT];
  for (int k = 0; k < NVERT; ++k) v[k] = vert[e * NVERT + k];
  // contract dimensions from the last to the first: after step d the first 2^d entries hold the partial results
  int len = NVERT;
  for (int d = ND - 1; d >= 0; --d) {
    const double x = c.node[(q / ipow(RS, ND - 1 - d)) % RS];
    len /= 2;
    for (int k = 0; k < len; ++k) {
      double acc = 0.;
      acc += (1. - x) * v[2 * k];
      acc += x * v[2 * k + 1];
      v[k] = acc;
    }
  }
  m.av[cls][e * 2 * NQ + NQ + q] = v[0];
}

}  // namespace hx
