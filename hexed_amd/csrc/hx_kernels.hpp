// gfx950 kernels of the DG hot path, FP64 form. Included once per n_dim by hx_kernels_nd*.hip, which
// instantiate them for row sizes 2..8.
//
// Element kernels run one thread per quadrature point and EPB = max(1, 256/nq) elements per workgroup; the
// per-element flux tensor, face values and new state are staged in LDS so that every derivative contraction
// and face extrapolation reads on-chip data. Connection kernels run one thread per face quadrature point.
#pragma once

#include <cfloat>
#include <type_traits>

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
  static constexpr int NRC = NV > RS ? NV : RS;         // residual-cache variables per point
  // LDS copies of per-point arrays carry one pad slot after every row along the last dimension, so that the
  // columns of that dimension (rows rs apart) fall on different banks
  static constexpr int NQP = NQ + NQ / RS;
  static __device__ int pad(int q) { return q + q / RS; }
};

template <int RS>
__device__ inline void load_ops(Ops<RS>& s, const Ops<RS>* g, int t, int nthreads)
{
  const double* src = reinterpret_cast<const double*>(g);
  double* dst = reinterpret_cast<double*>(&s);
  for (int i = t; i < int(sizeof(Ops<RS>) / sizeof(double)); i += nthreads) dst[i] = src[i];
}

__device__ inline double* slot_ptr(const Dev_mesh& m, int slot) { return m.faces + size_t(slot) * m.face_sz; }

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

// ---- element kernels: row contractions ------------------------------------------------------------------
// The element kernels alternate between point phases (one thread per quadrature point) and row phases, where one
// thread owns a "column" (an element, a variable and one row along a reference direction) and computes the whole
// row of outputs from the row's rs values and its two face values. The operator constants are a by-value kernel
// argument: the fully unrolled contractions read them with uniform scalar loads, so one LDS read of the data feeds
// rs multiply-adds.
// Arithmetic forms of the element kernels. FP64: the reference's arithmetic. FP32-MFMA: volume fluxes, gradients
// and derivative contractions in FP32, the contractions on the matrix cores with FP32 accumulation; the state, the
// RK update and residual cache, the face fluxes and their jumps against the interior extrapolation, the Jacobian
// determinant scaling and the face-buffer extrapolations stay FP64.
struct Fp64_form
{
  using T = double;
  static constexpr bool mfma = false;
};
struct Fp32_mfma_form
{
  using T = float;
  static constexpr bool mfma = true;
};

// The arithmetic form of the Euler Local for form P: the FP32-MFMA form runs the FP64 kernel on deformed elements, where
// its FP64 pressure rows (below) cost more LDS traffic and FP64 work than the matrix cores save (measured on the 48^3
// deformed mesh: 1.82 against 1.71 ms per launch)
template <class P, bool DEF>
using euler_form = std::conditional_t<DEF, Fp64_form, P>;

// Threads per workgroup: one per quadrature point, rounded up to whole wavefronts for the matrix-core form so that
// every lane of a wavefront takes part in its matrix instructions.
template <int ND, int RS, class P>
constexpr int form_block()
{
  constexpr int b = Shape<ND, RS>::BLOCK;
  return P::mfma ? (b + 63) / 64 * 64 : b;
}

// Calls f(ok, el, v, r, base, sd) for every column (el, v, r) of direction i_dim: v < NVAR, r the row counter (face
// qpoint), base the row's first point and sd its stride in the padded LDS layout (Row_index.hpp:14-64). With
// UNIFORM, every thread takes part in each round (ok is false past the last column and the indices are then those of
// column 0).
template <int ND, int RS, int NVAR, int NT, bool UNIFORM, class F>
__device__ inline void for_columns(int i_dim, F&& f)
{
  using S = Shape<ND, RS>;
  using R = Row<ND, RS>;
  constexpr int PER = NVAR * S::NFQ, TOTAL = S::EPB * PER;
  const int sd = R::stride(i_dim);
  const int sdp = i_dim == ND - 1 ? 1 : sd + sd / RS;
  if constexpr (UNIFORM) {
    for (int c0 = 0; c0 < TOTAL; c0 += NT) {
      const int c = c0 + int(threadIdx.x);
      const bool ok = c < TOTAL;
      const int cc = ok ? c : 0;
      const int el = cc / PER, vr = cc % PER, v = vr / S::NFQ, r = vr % S::NFQ;
      f(ok, el, v, r, S::pad(R::base_of_fq(r, i_dim)), sdp);
    }
  } else {
    for (int c = threadIdx.x; c < TOTAL; c += NT) {
      const int el = c / PER, vr = c % PER, v = vr / S::NFQ, r = vr % S::NFQ;
      f(true, el, v, r, S::pad(R::base_of_fq(r, i_dim)), sdp);
    }
  }
}

template <int RS, class T>
__device__ inline void load_row(const T* p, int sd, T (&x)[RS])
{
#pragma unroll
  for (int k = 0; k < RS; ++k) x[k] = p[k * sd];
}

// y = M x for an rs x rs matrix M
template <int RS, class T>
__device__ inline void row_matvec(const T (&mat)[RS][RS], const T (&x)[RS], T (&y)[RS])
{
#pragma unroll
  for (int i = 0; i < RS; ++i) {
    T a = 0;
#pragma unroll
    for (int k = 0; k < RS; ++k) a += mat[i][k] * x[k];
    y[i] = a;
  }
}

// out[base + i sd] = -y[i] (first direction), else out -= y
template <int RS, class T>
__device__ inline void sub_row(T* out, int sd, const T (&y)[RS], bool first)
{
#pragma unroll
  for (int i = 0; i < RS; ++i) out[i * sd] = first ? -y[i] : out[i * sd] - y[i];
}

// boundary(s) . x in FP64 (Spatial.hpp:29-71)
template <int RS, class T, class O>
__device__ inline double row_bound(const O& o, int s, const T (&x)[RS])
{
  double a = 0;
#pragma unroll
  for (int k = 0; k < RS; ++k) a += o.bound[s][k] * double(x[k]);
  return a;
}

// The operator constants read through a constant-address-space pointer that the compiler cannot follow across the
// empty asm: each call is reloaded with scalar loads where it is used, instead of being held (and spilled to vector
// register lanes) in scalar registers across a kernel's phases.
template <int RS>
struct Ops_reload
{
  const Eops<RS, double>* g;
  __device__ const auto& operator()() const
  {
    auto r = (const __attribute__((address_space(4))) Eops<RS, double>*)g;
    asm volatile("" : "+s"(r));
    return *r;
  }
};

// Row derivatives, Derivative.hpp:51-60: faces(x, b0, b1) = D x + L (b - B x), interior(x) = D x.
template <int RS, class P, bool RELOAD = false>
struct Contraction;

// FP64: the reference's operation order, D x + L (b - B x), on the vector units. (Folding the lift into
// (D - L B) x + L b saves 2 rs multiply-adds per column but rounds the two large face terms separately, which
// measurably loosens agreement with the reference arithmetic.) `bx` returns B x, the row's face extrapolation.
// For even row sizes the Gauss-Legendre tables are exactly (anti)symmetric -- D[rs-1-i][k] = -D[i][rs-1-k],
// B[1][k] = B[0][rs-1-k], L[i][1] = -L[rs-1-i][0], checked bit for bit by the CPU tests -- so the mirrored rows take
// their coefficients from the first half with a sign modifier: the same products in the same order, from half as
// many scalar registers.
template <int RS, bool RELOAD>
struct Contraction<RS, Fp64_form, RELOAD>
{
  typedef const __attribute__((address_space(4))) Eops<RS, double>* cptr;
  const Eops<RS, double>& o;   // by-value kernel argument
  const Eops<RS, double>* g;   // device copy, read when RELOAD
  __device__ Contraction(const Eops<RS, double>& o_, const Eops<RS, double>* dev) : o(o_), g(dev) {}
  // The constants come either from the by-value kernel argument, which the compiler loads into scalar registers
  // once, or with RELOAD (for kernels whose other kernel arguments already fill the scalar registers) through a
  // constant-address-space pointer that the compiler cannot follow across the empty asm, one row (pair) at a time
  // behind a scheduling barrier, so that each group is reloaded with scalar loads when it is needed.
  template <class Get>
  __device__ void run(Get&& get, bool with_faces, const double (&x)[RS], double b0, double b1, double (&y)[RS],
                      double (&bx)[2]) const
  {
    {
      const auto& c = get();
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        double a = 0;
#pragma unroll
        for (int k = 0; k < RS; ++k) {
          if constexpr (RS % 2 == 0) a += (s == 0 ? c.bound[0][k] : c.bound[0][RS - 1 - k]) * x[k];
          else a += c.bound[s][k] * x[k];
        }
        bx[s] = a;
      }
    }
    const double j0 = with_faces ? b0 - bx[0] : 0., j1 = with_faces ? b1 - bx[1] : 0.;
    if constexpr (RS % 2 == 0) {
#pragma unroll
      for (int i = 0; i < RS / 2; ++i) {
        if constexpr (RELOAD) __builtin_amdgcn_sched_barrier(0);
        const auto& c = get();
        const int im = RS - 1 - i;
        double a = 0, am = 0;
#pragma unroll
        for (int k = 0; k < RS; ++k) {
          a += c.diff[i][k] * x[k];
          am += -c.diff[i][RS - 1 - k] * x[k];
        }
        const double li = c.lift[i][0], lm = c.lift[im][0];
        y[i] = with_faces ? a + (li * j0 + (-lm) * j1) : a;
        y[im] = with_faces ? am + (lm * j0 + (-li) * j1) : am;
      }
    } else {
#pragma unroll
      for (int i = 0; i < RS; ++i) {
        if constexpr (RELOAD) __builtin_amdgcn_sched_barrier(0);
        const auto& c = get();
        double a = 0;
#pragma unroll
        for (int k = 0; k < RS; ++k) a += c.diff[i][k] * x[k];
        y[i] = with_faces ? a + (c.lift[i][0] * j0 + c.lift[i][1] * j1) : a;
      }
    }
    if constexpr (RELOAD) __builtin_amdgcn_sched_barrier(0);
  }
  // D x + L (b - B x) where `with_faces`, else D x
  __device__ void faces_if(bool with_faces, const double (&x)[RS], double b0, double b1, double (&y)[RS],
                           double (&bx)[2]) const
  {
    if constexpr (RELOAD) {
      run([&]() -> const auto& {
            cptr r = (cptr)g;
            asm volatile("" : "+s"(r));
            return *r;
          },
          with_faces, x, b0, b1, y, bx);
    } else {
      run([&]() -> const Eops<RS, double>& { return o; }, with_faces, x, b0, b1, y, bx);
    }
  }
  __device__ void faces(const double (&x)[RS], double b0, double b1, double (&y)[RS]) const
  {
    double bx[2];
    faces_if(true, x, b0, b1, y, bx);
  }
};

// FP32-MFMA: V_MFMA_F32_4X4X1_16B_F32 computes 16 independent 4x4 outer products per wavefront, lane 4b + j holding
// column j of block b. Each lane owns one column of data (its B operand) and row 4t + (lane mod 4) of the stacked
// operator [D; B] (its A operand), so after rs rank-1 steps lane 4b + j holds D x and B x of its own row. The face
// jumps b - B x are formed in FP64 and lifted by two more rank-1 steps with the rows of L.
typedef float hx_f4 __attribute__((ext_vector_type(4)));

template <int RS, bool RELOAD>
struct Contraction<RS, Fp32_mfma_form, RELOAD>
{
  static constexpr int TA = (RS + 2 + 3) / 4;  // tiles of the stacked [D; B] rows
  static constexpr int TL = (RS + 3) / 4;      // tiles holding rows of D (the lifted ones)
  float ax[TA][RS];
  float aj[TL][2];
  __device__ Contraction(const Eops<RS, double>& o, const Eops<RS, double>*)
  {
    const int lane_row = threadIdx.x & 3;
#pragma unroll
    for (int t = 0; t < TA; ++t)
#pragma unroll
      for (int k = 0; k < RS; ++k) {
        float val = 0.f;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const int row = 4 * t + rr;
          if (lane_row == rr) {
            if (row < RS) val = float(o.diff[row][k]);
            else if (row < RS + 2) val = float(o.bound[row - RS][k]);
          }
        }
        ax[t][k] = val;
      }
#pragma unroll
    for (int t = 0; t < TL; ++t)
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        float val = 0.f;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr)
          if (lane_row == rr && 4 * t + rr < RS) val = float(o.lift[4 * t + rr][s]);
        aj[t][s] = val;
      }
  }
  // D x + L (b - B x) where `with_faces`, else D x; without divergence around the matrix instructions (lanes
  // without faces lift zero jumps)
  __device__ void faces_if(bool with_faces, const float (&x)[RS], double b0, double b1, float (&y)[RS],
                           float (&bx)[2]) const
  {
    hx_f4 acc[TA];
#pragma unroll
    for (int t = 0; t < TA; ++t) acc[t] = hx_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < RS; ++k)
#pragma unroll
      for (int t = 0; t < TA; ++t) acc[t] = __builtin_amdgcn_mfma_f32_4x4x1f32(ax[t][k], x[k], acc[t], 0, 0, 0);
    bx[0] = acc[RS / 4][RS % 4];
    bx[1] = acc[(RS + 1) / 4][(RS + 1) % 4];
    const float j0 = with_faces ? float(b0 - double(bx[0])) : 0.f;
    const float j1 = with_faces ? float(b1 - double(bx[1])) : 0.f;
#pragma unroll
    for (int t = 0; t < TL; ++t) {
      acc[t] = __builtin_amdgcn_mfma_f32_4x4x1f32(aj[t][0], j0, acc[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_4x4x1f32(aj[t][1], j1, acc[t], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < RS; ++i) y[i] = acc[i / 4][i % 4];
  }
  __device__ void faces(const float (&x)[RS], double b0, double b1, float (&y)[RS]) const
  {
    float bx[2];
    faces_if(true, x, b0, b1, y, bx);
  }
};

// The same contraction with the operands read from a workgroup copy of Eops::mfma in LDS (lds_mfma_operands) at every
// call, instead of being formed in registers from the FP64 constants: each lane reads its lane row's tile with 16-byte
// reads, which frees the scalar registers the constants took and the conversions and selects that formed them. The
// values, products and their order are the ones of Contraction<RS, Fp32_mfma_form>.
template <int RS>
struct Contraction_lds
{
  using E = Eops<RS, double>;
  static constexpr int TA = E::MF_TA, TL = E::MF_TL;
  const float* row;  // this lane's row of the LDS copy
  __device__ explicit Contraction_lds(const float* tab) : row(tab + (threadIdx.x & 3) * E::MF_STRIDE) {}
  __device__ void faces_if(bool with_faces, const float (&x)[RS], double b0, double b1, float (&y)[RS],
                           float (&bx)[2]) const
  {
    float a[E::MF_STRIDE];
    const hx_f4* r4 = reinterpret_cast<const hx_f4*>(row);
#pragma unroll
    for (int c = 0; c < E::MF_STRIDE / 4; ++c) {
      const hx_f4 v = r4[c];
      a[4 * c] = v[0];
      a[4 * c + 1] = v[1];
      a[4 * c + 2] = v[2];
      a[4 * c + 3] = v[3];
    }
    hx_f4 acc[TA];
#pragma unroll
    for (int t = 0; t < TA; ++t) acc[t] = hx_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < RS; ++k)
#pragma unroll
      for (int t = 0; t < TA; ++t) acc[t] = __builtin_amdgcn_mfma_f32_4x4x1f32(a[t * RS + k], x[k], acc[t], 0, 0, 0);
    bx[0] = acc[RS / 4][RS % 4];
    bx[1] = acc[(RS + 1) / 4][(RS + 1) % 4];
    const float j0 = with_faces ? float(b0 - double(bx[0])) : 0.f;
    const float j1 = with_faces ? float(b1 - double(bx[1])) : 0.f;
#pragma unroll
    for (int t = 0; t < TL; ++t) {
      acc[t] = __builtin_amdgcn_mfma_f32_4x4x1f32(a[TA * RS + 2 * t], j0, acc[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_4x4x1f32(a[TA * RS + 2 * t + 1], j1, acc[t], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < RS; ++i) y[i] = acc[i / 4][i % 4];
  }
  __device__ void faces(const float (&x)[RS], double b0, double b1, float (&y)[RS]) const
  {
    float bx[2];
    faces_if(true, x, b0, b1, y, bx);
  }
};

// Copies the device operator table's matrix-core operands into the workgroup's LDS (4 * Eops::MF_STRIDE floats at
// tab); the caller synchronizes before the first contraction
template <int RS, int NT>
__device__ inline void lds_mfma_operands(const Eops<RS, double>* od, float* tab)
{
  using E = Eops<RS, double>;
  const float* src = &od->mfma[0][0];
  for (int k = threadIdx.x; k < 4 * E::MF_STRIDE; k += NT) tab[k] = src[k];
}

// Modal filter of NVAR padded [NQP] arrays per element at `buf` (element stride `per`), in place, one direction at a
// time (Spatial.hpp:472-481). Ends synchronized.
template <int ND, int RS, int NVAR, int NT, class T>
__device__ inline void filter_columns(const Eops<RS, double>& o, T* buf, int per)
{
  using S = Shape<ND, RS>;
  T filt[RS][RS];
  for (int i = 0; i < RS; ++i)
    for (int k = 0; k < RS; ++k) filt[i][k] = T(o.filt[i][k]);
  for (int i = 0; i < ND; ++i) {
    __syncthreads();
    for_columns<ND, RS, NVAR, NT, false>(i, [&](bool, int el, int v, int, int base, int sd) {
      T* row = buf + el * per + v * S::NQP + base;
      T x[RS], y[RS];
      load_row(row, sd, x);
      row_matvec(filt, x, y);
#pragma unroll
      for (int k = 0; k < RS; ++k) row[k * sd] = y[k];
    });
  }
  __syncthreads();
}

// Loads block `ldg` (0: state/flux, 1: LDG) of every face of the workgroup's elements into dst[EPB][2ND][NV][NFQ]
// (element stride `per`).
template <int ND, int RS, int NT>
__device__ inline void load_faces(const Dev_mesh& m, int cls, int blk, int ldg, double* dst, int per)
{
  using S = Shape<ND, RS>;
  constexpr int PF = 2 * ND * S::NV * S::NFQ;
  const int n = m.n_elem[cls];
  for (int k = threadIdx.x; k < S::EPB * PF; k += NT) {
    const int el = k / PF, r = k % PF, f = r / (S::NV * S::NFQ), vq = r % (S::NV * S::NFQ);
    const int e = blk * S::EPB + el;
    if (e < n) {
      const int slot = m.face_slot[cls][e * 2 * ND + f];
      dst[el * per + r] = (ldg ? slot_ptr(m, slot) + S::NV * S::NFQ : state_ptr(m, slot, m.alt_rd))[vq];
    } else {
      dst[el * per + r] = 0.;
    }
  }
}

// Write_face of the FP64 padded [NVAR][NQP] arrays at `buf` (element stride `per`) into block 0 of every face
// (Spatial.hpp:29-71)
template <int ND, int RS, int NT, int NVAR = Shape<ND, RS>::NV>
__device__ inline void write_faces_from(const Dev_mesh& m, int cls, int blk, const Eops<RS, double>& o, const double* buf,
                                        int per)
{
  using S = Shape<ND, RS>;
  const int n = m.n_elem[cls];
  for (int i = 0; i < ND; ++i) {
    for_columns<ND, RS, NVAR, NT, false>(i, [&](bool, int el, int v, int r, int base, int sd) {
      const int e = blk * S::EPB + el;
      if (e >= n) return;
      double x[RS];
      load_row(buf + el * per + v * S::NQP + base, sd, x);
      const int* fs = m.face_slot[cls] + e * 2 * ND + 2 * i;
      state_ptr(m, fs[0], m.alt_wr)[v * S::NFQ + r] = row_bound(o, 0, x);
      state_ptr(m, fs[1], m.alt_wr)[v * S::NFQ + r] = row_bound(o, 1, x);
    });
  }
}

// ---- Neighbor (Spatial.hpp:613-704) ----------------------------------------------------------------------
// The numerical flux at face point f of connection con of class DEF, and where each side keeps it: side 0 at f, side
// 1 at p1, each times its normal sign. With VISC the LDG average of the two states goes into both LDG blocks.
template <int ND, int RS>
struct Con_flux
{
  double fl[ND + 2];
  double* f0;
  double* f1;
  int s0, p1;
  double sg0, sg1;
};

template <int ND, int RS, bool DEF, bool VISC>
__device__ __forceinline__ Con_flux<ND, RS> con_flux(const Dev_mesh& m, int con, int f)
{
  constexpr int NFQ = Shape<ND, RS>::NFQ, NV = ND + 2;
  constexpr int cls = DEF ? HX_DEFORMED : HX_CARTESIAN;
  Con_flux<ND, RS> r;
  const Dir dir(m.con_dir[cls][con]);
  const int s0 = m.con_slot[cls][2 * con], s1 = m.con_slot[cls][2 * con + 1];
  r.s0 = s0;
  r.f0 = state_ptr(m, s0, m.alt_rd);
  r.f1 = state_ptr(m, s1, m.alt_rd);
  r.p1 = f;
  r.sg0 = r.sg1 = 1.;
  double n[ND];
  if constexpr (DEF) {
    r.p1 = face1_index<ND, RS>(dir, f);
    r.sg0 = 1 - 2 * int(dir.flip_normal(0));
    r.sg1 = 1 - 2 * int(dir.flip_normal(1));
    const double* nn = m.normals + size_t(m.con_nslot[2 * con]) * ND * NFQ;
    for (int j = 0; j < ND; ++j) n[j] = r.sg0 * nn[j * NFQ + f];
  } else {
    for (int j = 0; j < ND; ++j) n[j] = (j == dir.i_dim0);
  }
  double u0[NV], u1[NV];
  for (int v = 0; v < NV; ++v) {
    u0[v] = r.f0[v * NFQ + f];
    u1[v] = r.f1[v * NFQ + r.p1];
  }
  if constexpr (VISC) {
    double* l0 = slot_ptr(m, s0) + NV * NFQ;
    double* l1 = slot_ptr(m, s1) + NV * NFQ;
    for (int v = 0; v < NV; ++v) {
      const double avg = .5 * (u0[v] + u1[v]);
      l0[v * NFQ + f] = avg;
      l1[v * NFQ + r.p1] = avg;
    }
  }
  double F0[NV], F1[NV];
  flux_conv<ND>(u0, n, F0);
  flux_conv<ND>(u1, n, F1);
  double nsq = 0;
  for (int j = 0; j < ND; ++j) nsq += n[j] * n[j];
  const double c = fmax(char_speed<ND>(u0), char_speed<ND>(u1)) * sqrt(nsq);
  for (int v = 0; v < NV; ++v) r.fl[v] = .5 * (F0[v] + F1[v] + c * (u0[v] - u1[v]));
  return r;
}

template <int ND, int RS, bool DEF, bool VISC>
__device__ __forceinline__ void neighbor_con(const Dev_mesh& m, int con, int f)
{
  constexpr int NFQ = Shape<ND, RS>::NFQ, NV = ND + 2;
  const Con_flux<ND, RS> r = con_flux<ND, RS, DEF, VISC>(m, con, f);
  for (int v = 0; v < NV; ++v) {
    r.f0[v * NFQ + f] = r.sg0 * r.fl[v];
    r.f1[v * NFQ + r.p1] = r.sg1 * r.fl[v];
  }
}

// one thread per (connection, face point) item `gid` of the class's connections
template <int ND, int RS, bool DEF, bool VISC>
__device__ __forceinline__ void neighbor_at(const Dev_mesh& m, int gid)
{
  constexpr int NFQ = Shape<ND, RS>::NFQ;
  const int con = gid / NFQ, f = gid % NFQ;
  if (con < m.n_con[DEF ? HX_DEFORMED : HX_CARTESIAN]) neighbor_con<ND, RS, DEF, VISC>(m, con, f);
}

template <int ND, int RS, bool DEF, bool VISC>
__global__ __launch_bounds__(256) void k_neighbor(Dev_mesh m)
{
  neighbor_at<ND, RS, DEF, VISC>(m, blockIdx.x * 256 + threadIdx.x);
}

// Neighbor over the connections that touch a hanging-face mortar (Dev_mesh::mortar_con), whose fluxes the face-flux
// folding leaves to a launch before the Local
template <int ND, int RS, bool DEF, bool VISC>
__global__ __launch_bounds__(256) void k_neighbor_mortar(Dev_mesh m)
{
  constexpr int NFQ = Shape<ND, RS>::NFQ;
  constexpr int cls = DEF ? HX_DEFORMED : HX_CARTESIAN;
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int i = gid / NFQ, f = gid % NFQ;
  if (i < m.n_mortar_con[cls]) neighbor_con<ND, RS, DEF, VISC>(m, m.mortar_con[cls][i], f);
}

// Face-flux folding: the numerical flux of face point p of face `face` of element e (class cls), times this side's
// sign, computed from the two extrapolated states exactly as Neighbor computes it for the connection (Spatial.hpp:
// 613-704, the operands in the connection's side order), and with VISC the LDG average of the two states. A face whose
// flux a launch before the Local computed (hanging faces) reads flux and average from its own slot. On a boundary
// connection the element also writes the ghost side's flux into the ghost slot's block 0 in the primary buffer, where
// Neighbor leaves it.
template <int ND, int RS, bool VISC>
__device__ inline void face_flux(const Dev_mesh& m, int cls, int e, int face, int p, double (&fl)[ND + 2],
                                 double (&avg)[ND + 2])
{
  constexpr int NFQ = Shape<ND, RS>::NFQ, NV = ND + 2;
  const size_t k = size_t(e) * 2 * ND + face;
  const int* pi = m.partner[cls] + 4 * k;
  const int own_slot = m.face_slot[cls][k];
  const double* own = state_ptr(m, own_slot, m.alt_rd);
  const int ps = pi[0];
  if (ps < 0) {
    for (int v = 0; v < NV; ++v) fl[v] = own[v * NFQ + p];
    if constexpr (VISC) {
      const double* l = slot_ptr(m, own_slot) + NV * NFQ;
      for (int v = 0; v < NV; ++v) avg[v] = l[v * NFQ + p];
    }
    return;
  }
  const int info = pi[1];
  const Dir dir(info & 63);
  const bool side1 = (info >> 6) & 1, def = (info >> 7) & 1, bc = (info >> 8) & 1;
  const double* par = state_ptr(m, ps, m.alt_rd);
  int f0 = p, f1 = p;
  double sg0 = 1., sg1 = 1.;
  double n[ND];
  if (def) {
    if (side1) f0 = face0_index<ND, RS>(dir, p);
    else f1 = face1_index<ND, RS>(dir, p);
    sg0 = 1 - 2 * int(dir.flip_normal(0));
    sg1 = 1 - 2 * int(dir.flip_normal(1));
    const double* nn = m.normals + size_t(pi[2]) * ND * NFQ;
    for (int j = 0; j < ND; ++j) n[j] = sg0 * nn[j * NFQ + f0];
  } else {
    for (int j = 0; j < ND; ++j) n[j] = (j == dir.i_dim0);
  }
  const double* a0 = side1 ? par : own;
  const double* a1 = side1 ? own : par;
  double u0[NV], u1[NV];
  for (int v = 0; v < NV; ++v) {
    u0[v] = a0[v * NFQ + f0];
    u1[v] = a1[v * NFQ + f1];
  }
  if constexpr (VISC)
    for (int v = 0; v < NV; ++v) avg[v] = .5 * (u0[v] + u1[v]);
  double F0[NV], F1[NV];
  flux_conv<ND>(u0, n, F0);
  flux_conv<ND>(u1, n, F1);
  double nsq = 0;
  for (int j = 0; j < ND; ++j) nsq += n[j] * n[j];
  const double c = fmax(char_speed<ND>(u0), char_speed<ND>(u1)) * sqrt(nsq);
  double* ghost = (bc && !side1) ? slot_ptr(m, ps) : nullptr;
  for (int v = 0; v < NV; ++v) {
    const double x = .5 * (F0[v] + F1[v] + c * (u0[v] - u1[v]));
    fl[v] = (side1 ? sg1 : sg0) * x;
    if (ghost) ghost[v * NFQ + f1] = sg1 * x;
  }
}

// The faces' numerical fluxes of the workgroup's elements into dst[EPB][2ND][NV][NFQ] (element stride `per`), one thread
// per face point (face-flux folding, in place of load_faces)
template <int ND, int RS, int NT>
__device__ inline void fold_fluxes(const Dev_mesh& m, int cls, int blk, double* dst, int per)
{
  using S = Shape<ND, RS>;
  constexpr int NFP = 2 * ND * S::NFQ, NV = S::NV, NFQ = S::NFQ;
  const int n = m.n_elem[cls];
  for (int k = threadIdx.x; k < S::EPB * NFP; k += NT) {
    const int el = k / NFP, r = k % NFP, face = r / NFQ, p = r % NFQ;
    const int e = blk * S::EPB + el;
    double fl[NV], avg[NV];
    if (e < n) face_flux<ND, RS, false>(m, cls, e, face, p, fl, avg);
    else
      for (int v = 0; v < NV; ++v) fl[v] = 0.;
    for (int v = 0; v < NV; ++v) dst[el * per + (face * NV + v) * NFQ + p] = fl[v];
  }
}

// ---- convective flux rows of the two forms ---------------------------------------------------------------------
// FP64: the flux row x of direction i enters -(D x + L (b - B x)) as is.
// FP32-MFMA: the FP32 operands carry only what FP32 holds accurately relative to the small residual of a nearly
// uniform flow:
//  * every flux row is shifted by its numerical face flux b0 before rounding; D 1 = 0 and B 1 = 1 make
//    D x + L (b - B x) invariant under x -> x - b0, b -> b - b0, so the FP32 operands carry only the variation
//    along the row;
//  * except the pressure part p n_iv of the momentum fluxes of deformed elements, which varies along a row as much as
//    the normal n_iv does, so that no shift makes it small. It stays FP64 (LDS row xp, array PD): its face
//    extrapolation B xp comes off the jumps of the rest of the row, which goes to the matrix cores unshifted, and
//    its derivative D xp is summed over the directions in FP64 (array RP). (Summing it in FP32 would round partial
//    sums many times larger than the total.)
template <int ND, bool DEF>
__device__ inline bool has_pressure(int v)
{
  return DEF && v < ND;
}

// point phase: this point's operands of direction i (normal nrm = n_i.), from the state u with pressure p; sb0 is
// the point's row's numerical flux at face 0 (variable stride nfq)
template <int ND, bool DEF, class P>
__device__ inline void conv_point(const double (&u)[ND + 2], double p, const double (&nrm)[ND], const double* sb0, int nfq,
                                  typename P::T* f, double* pd, int stride)
{
  using T = typename P::T;
  double fl[ND + 2];
  flux_conv<ND>(u, nrm, fl);
#pragma unroll
  for (int v = 0; v < ND + 2; ++v) {
    if constexpr (!P::mfma) {
      f[v * stride] = fl[v];
    } else if (has_pressure<ND, DEF>(v)) {
      const double xp = p * nrm[v];
      pd[v * stride] = xp;
      f[v * stride] = T(fl[v] - xp);
    } else {
      f[v * stride] = T(fl[v] - sb0[v * nfq]);
    }
  }
}

// row phase: y = D x + L (b - B x) of the flux row of variable v from the operands conv_point wrote (frow; prow the
// pressure row), without the FP32-MFMA form's FP64 pressure derivative (pressure_row). The matrix instructions run
// outside any branch that diverges by variable.
template <int ND, int RS, bool DEF, class P, class C, class G>
__device__ inline void conv_row(const C& con, G&& ops, int v, const typename P::T* frow, const double* prow, int sd,
                                double b0, double b1, typename P::T (&y)[RS])
{
  using T = typename P::T;
  T x[RS];
  load_row(frow, sd, x);
  if constexpr (!P::mfma) {
    con.faces(x, b0, b1, y);
  } else {
    double s0 = 0., s1 = b1 - b0;
    if (has_pressure<ND, DEF>(v)) {
      double xp[RS];
      load_row(prow, sd, xp);
      const auto& o = ops();
      s0 = b0 - row_bound(o, 0, xp);
      s1 = b1 - row_bound(o, 1, xp);
    }
    con.faces(x, s0, s1, y);
  }
}

// FP32-MFMA form on deformed elements: acc = D xp (first direction) or acc += D xp, in FP64, for a momentum row. The
// operator comes from `ops()` one row at a time (behind a scheduling barrier where `ops` reloads it).
template <int ND, int RS, bool DEF, class P, class G>
__device__ inline void pressure_row(G&& ops, int i, int v, const double* prow, double* acc, int sd)
{
  if constexpr (P::mfma && DEF) {
    if (v < ND) {
      double xp[RS];
      load_row(prow, sd, xp);
#pragma unroll
      for (int r = 0; r < RS; ++r) {
        if constexpr (std::is_same_v<std::decay_t<G>, Ops_reload<RS>>) __builtin_amdgcn_sched_barrier(0);
        const auto& o = ops();
        double a = 0;
#pragma unroll
        for (int k = 0; k < RS; ++k) a += o.diff[r][k] * xp[k];
        acc[r * sd] = i == 0 ? a : acc[r * sd] + a;
      }
    }
  }
}

// 1/time step at one quadrature point (Spatial.hpp:790-818): the convective scale over the interpolated vertex spacing,
// plus the diffusive one for Navier-Stokes
template <int ND, int RS, bool VISC>
__device__ inline double max_dt_scale_u(const Dev_mesh& m, const Ops<RS>* __restrict__ gops, const Ns_params& ns, int cls,
                                        size_t e, int q, const double (&u)[ND + 2], double max_cfl_c, double max_cfl_d)
{
  constexpr int NQ = Shape<ND, RS>::NQ, NVERT = 1 << ND;
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
  double scale = char_speed<ND>(u) / max_cfl_c / spacing;
  if constexpr (VISC) {
    double bulk = 0., lap = 0.;
    if (m.av[cls]) {
      bulk = m.av[cls][e * 2 * NQ + q];
      lap = m.av[cls][e * 2 * NQ + NQ + q];
    }
    scale += diffusivity<ND>(ns, u, bulk, lap) / max_cfl_d / spacing / spacing;
  }
  return scale;
}

// The stage coefficient: `coef` itself, or with a device-resident time step (dtp not null) coef times that time step,
// the product the host forms when it passes the time step by value (coef = dt, or dt times the stage-1 step ratio)
__device__ inline double step_coef(double coef, const double* dtp) { return dtp ? coef * *dtp : coef; }

// ---- Local, Euler (Spatial.hpp:326-509 with Navier_stokes<false>) -----------------------------------------
template <int ND, int RS, bool DEF, class P>
struct Euler_lds
{
  using S = Shape<ND, RS>;
  // arrays of T (padded)
  static constexpr int F = 0;                               // [NV][NQP] flux along the current direction
  static constexpr int RT = F + S::NV * S::NQP;             // [NV][NQP] time rate
  static constexpr int PER_T = RT + S::NV * S::NQP;
  // arrays of double
  static constexpr int SB = 0;                              // [2ND][NV][NFQ] numerical face fluxes
  static constexpr int PD = SB + 2 * ND * S::NV * S::NFQ;   // [ND][NQP] FP32-MFMA, deformed: pressure part p n_iv
  static constexpr int RP = PD + ND * S::NQP;               // [ND][NQP] FP32-MFMA, deformed: its derivative, summed
  static constexpr int LOOP_D = PD + (P::mfma && DEF ? 2 * ND * S::NQP : 0);
  static constexpr int U = 0;                               // [NV][NQP] updated state, for Write_face (after the loop)
  static constexpr int PER_D = LOOP_D > S::NV * S::NQP ? LOOP_D : S::NV * S::NQP;
};

// The Euler Local's workgroup: whole wavefronts in both forms (the fused time-step reduction shuffles across full
// wavefronts)
template <int ND, int RS>
constexpr int euler_block()
{
  return (Shape<ND, RS>::BLOCK + 63) / 64 * 64;
}

// Waves per SIMD that the workgroup's LDS allows (160 KB per CU, 4 SIMDs), at most 5: the register target given to
// the compiler, so that registers do not lower the occupancy below what LDS permits. With face-flux folding, where a
// thread computes more than one face point's flux, at most 4 (the registers of two face points' states and fluxes
// would spill at 5).
template <int ND, int RS, bool DEF, class P, bool FOLD = false>
constexpr int euler_waves()
{
  using L = Euler_lds<ND, RS, DEF, P>;
  constexpr int lds = Shape<ND, RS>::EPB * (int(sizeof(typename P::T)) * L::PER_T + 8 * L::PER_D);
  constexpr int w = (163840 / lds) * (euler_block<ND, RS>() / 64) / 4;
  constexpr int cap = FOLD && Shape<ND, RS>::EPB * 2 * ND * Shape<ND, RS>::NFQ > euler_block<ND, RS>() ? 4 : 5;
  return w < 1 ? 1 : (w > cap ? cap : w);
}

// LDS bytes of one Euler Local workgroup: the arrays of T, then (8-byte aligned) the arrays of double
template <int ND, int RS, bool DEF, class P>
constexpr int euler_lds_bytes()
{
  using L = Euler_lds<ND, RS, DEF, P>;
  constexpr int t = Shape<ND, RS>::EPB * int(sizeof(typename P::T)) * L::PER_T;
  return (t + 7) / 8 * 8 + Shape<ND, RS>::EPB * 8 * L::PER_D;
}

// One workgroup's worth of the Euler Local: the EPB elements of tile `blk` of class DEF, with the LDS arrays lt (T) and
// ld (double) of Euler_lds. Shared by k_local_euler (one tile per workgroup) and the fused launch (tiles of several
// meshes, classes and row sizes per kernel).
// FOLD: the faces' numerical fluxes are computed here from both sides' states (face_flux) instead of read from the slots
// mtab: the FP32-MFMA form's LDS copy of the matrix-core operands (4 * Eops::MF_STRIDE floats, 16-byte aligned)
template <int ND, int RS, bool DEF, class P, bool FOLD = false>
__device__ __forceinline__ void local_euler_tile(const Dev_mesh& m, int blk, typename P::T* lt, double* ld, float* mtab,
                                                 const Eops<RS, double>& o, const Eops<RS, double>* od, double coef,
                                                 const double* __restrict__ dtp, int stage, int compute_residual,
                                                 int use_filter, const Dt_fuse& fuse, const Ops<RS>* __restrict__ gops)
{
  using S = Shape<ND, RS>;
  using L = Euler_lds<ND, RS, DEF, P>;
  using T = typename P::T;
  constexpr int NQ = S::NQ, NQP = S::NQP, NFQ = S::NFQ, NV = S::NV, EPB = S::EPB, NT = euler_block<ND, RS>();
  constexpr int cls = DEF ? HX_DEFORMED : HX_CARTESIAN;
  const int t = threadIdx.x;
  const bool pt = t < EPB * NQ;
  const int el = pt ? t / NQ : 0, q = pt ? t % NQ : 0, qp = S::pad(q);
  const int n = m.n_elem[cls];
  const int e = blk * EPB + el;
  const bool valid = pt && e < n;
  T* my = lt + el * L::PER_T;
  // (the first contraction follows a barrier)
  if constexpr (P::mfma) lds_mfma_operands<RS, NT>(od, mtab);
  const auto con = [&] {
    if constexpr (P::mfma) return Contraction_lds<RS>(mtab);
    else return Contraction<RS, P>(o, od);
  }();

  double u[NV];
  const double* st = m.state[cls] + size_t(e) * NV * NQ;
  for (int v = 0; v < NV; ++v) u[v] = valid ? st[v * NQ + q] : (v >= ND ? 1. : 0.);
  // what the stage update at the end reads from memory -- the residual cache (stage 1), the nominal size, the time
  // step scale and (deformed) the Jacobian determinant. Folded kernels up to row size 6 load the last two at the start
  // of the last direction, so that their latency overlaps that direction's phases; on Cartesian elements the first two
  // as well (Euler Local 0.82 -> 0.78 ms per step on the 32^3 box, 0.48 -> 0.45 on the 16^3 tree). On deformed
  // elements those are read at the update: earlier, they cost more registers than their latency saves (1.78 -> 1.84 ms
  // on the 48^3 deformed mesh).
  constexpr bool EARLY = !DEF && FOLD && RS <= 6;
  double rcv[NV];
  double tss_q = 0., nom = 1., det_q = 1.;
  const auto load_update_operands = [&] {
    const double* rc = m.rcache[cls] + size_t(e) * S::NRC * NQ;
    for (int v = 0; v < NV; ++v) rcv[v] = (valid && stage) ? rc[v * NQ + q] : 0.;
    if (valid) nom = m.nom_size[cls][e];
  };
  const auto load_scale = [&] {
    if (valid) {
      tss_q = m.tss[cls][size_t(e) * NQ + q];
      if constexpr (DEF) det_q = m.metrics[size_t(e) * (ND * ND + 1) * NQ + ND * ND * NQ + q];
    }
  };
  // deformed: each direction's reference-level normals are loaded in that direction's point phase, so that they
  // are not held in registers across the row phases
  const double* mt = m.metrics + (DEF ? size_t(e) * (ND * ND + 1) * NQ + q : 0);
  const double p64 = P::mfma ? pressure_of<ND>(u) : 0.;
  if constexpr (FOLD) fold_fluxes<ND, RS, NT>(m, cls, blk, ld + L::SB, L::PER_D);
  else load_faces<ND, RS, NT>(m, cls, blk, 0, ld + L::SB, L::PER_D);
  for (int i = 0; i < ND; ++i) {
    if (i == ND - 1) {
      if constexpr (EARLY) load_update_operands();
      if constexpr (FOLD && RS <= 6) load_scale();
    }
    // volume flux along reference direction i (Spatial.hpp:404-419), in FP64
    double nrm[ND];
    for (int j = 0; j < ND; ++j) {
      if constexpr (DEF) nrm[j] = valid ? mt[(i * ND + j) * NQ] : double(i == j);
      else nrm[j] = double(i == j);
    }
    if (i || P::mfma) __syncthreads();
    if (pt)
      conv_point<ND, DEF, P>(u, p64, nrm, ld + el * L::PER_D + L::SB + (2 * i) * NV * NFQ + Row<ND, RS>::fq(q, i), NFQ,
                             my + L::F + qp, ld + el * L::PER_D + L::PD + qp, NQP);
    __syncthreads();
    // -derivative(row_f, face_f) (Spatial.hpp:444-457)
    for_columns<ND, RS, NV, NT, P::mfma>(i, [&](bool ok, int el2, int v, int r, int base, int sd) {
      T* b = lt + el2 * L::PER_T;
      double* db = ld + el2 * L::PER_D;
      T y[RS];
      conv_row<ND, RS, DEF, P>(con, [&]() -> const Eops<RS, double>& { return o; }, v, b + L::F + v * NQP + base, db + L::PD + (v % ND) * NQP + base, sd,
                               db[L::SB + ((2 * i) * NV + v) * NFQ + r], db[L::SB + ((2 * i + 1) * NV + v) * NFQ + r], y);
      if (ok) {
        sub_row(b + L::RT + v * NQP + base, sd, y, i == 0);
        pressure_row<ND, RS, DEF, P>([&]() -> const Eops<RS, double>& { return o; }, i, v, db + L::PD + (v % ND) * NQP + base, db + L::RP + (v % ND) * NQP + base, sd);
      }
    });
  }
  if (use_filter) {
    filter_columns<ND, RS, NV, NT>(o, lt + L::RT, L::PER_T);
    if constexpr (P::mfma && DEF) filter_columns<ND, RS, ND, NT>(o, ld + L::RP, L::PER_D);
  } else {
    __syncthreads();
  }
  // stage update (Spatial.hpp:483-503)
  if constexpr (!EARLY) load_update_operands();
  if constexpr (!(FOLD && RS <= 6)) load_scale();
  if (valid) {
    double mult = step_coef(coef, dtp) * tss_q / nom;
    if constexpr (DEF) mult /= det_q;
    double* rcw = m.rcache[cls] + size_t(e) * S::NRC * NQ;
    double* sw = m.state[cls] + size_t(e) * NV * NQ;
    for (int v = 0; v < NV; ++v) {
      double r = double(my[L::RT + v * NQP + qp]);
      if (has_pressure<ND, DEF>(v) && P::mfma) r -= ld[el * L::PER_D + L::RP + v * NQP + qp];
      if (stage) r -= rcv[v];
      else rcw[v * NQ + q] = r;
      r *= mult;
      if (compute_residual) rcw[v * NQ + q] = r;
      else u[v] += r;
    }
    if (!compute_residual)
      for (int v = 0; v < NV; ++v) sw[v * NQ + q] = u[v];
  }
  // the next time step's Max_dt (Spatial.hpp:790-818, global time step) on the state just written: each point's
  // 1/scale here (its loads then overlap the face writes below), the workgroup minimum after them
  double c = DBL_MAX;
  if ((fuse.block_min || fuse.point_scale[cls]) && valid && !compute_residual) {
    const double scale = fuse.visc ? max_dt_scale_u<ND, RS, true>(m, gops, fuse.ns, cls, e, q, u, fuse.max_cfl_c, fuse.max_cfl_d)
                                   : max_dt_scale_u<ND, RS, false>(m, gops, fuse.ns, cls, e, q, u, fuse.max_cfl_c, fuse.max_cfl_d);
    // folded as k_max_dt does: a NaN of this point never wins (std::min(dt, NaN) keeps dt in the reference)
    const double inv = 1. / scale;
    c = (inv < c) ? inv : c;
    if (fuse.point_scale[cls]) fuse.point_scale[cls][size_t(e) * NQ + q] = inv;
  }
  // Write_face of the updated state (Spatial.hpp:507)
  if (pt)
    for (int v = 0; v < NV; ++v) ld[el * L::PER_D + L::U + v * NQP + qp] = u[v];
  __syncthreads();
  write_faces_from<ND, RS, NT>(m, cls, blk, o, ld + L::U, L::PER_D);
  if (fuse.block_min) {
    for (int off = 32; off > 0; off /= 2) {
      const double x = __shfl_xor(c, off, 64);
      c = (x < c) ? x : c;
    }
    // the per-wavefront minima reuse the start of the Write_face staging array, once every thread has read it (no LDS
    // of its own: at 3D row size 4 in FP64 that keeps four workgroups per compute unit)
    __syncthreads();
    double* wmin = ld;
    if ((t & 63) == 0) wmin[t / 64] = c;
    __syncthreads();
    if (t == 0) {
      double b = wmin[0];
      for (int w = 1; w < NT / 64; ++w) b = (wmin[w] < b) ? wmin[w] : b;
      fuse.block_min[fuse.block_offset + blk] = b;
    }
  }
}

template <int ND, int RS, bool DEF, class P, bool FOLD>
__global__ __launch_bounds__((euler_block<ND, RS>())) __attribute__((amdgpu_waves_per_eu(euler_waves<ND, RS, DEF, P, FOLD>())))
void k_local_euler(Dev_mesh m, Eops<RS, double> o, const Eops<RS, double>* od, double coef, const double* __restrict__ dtp,
                   int stage, int compute_residual, int use_filter, Dt_fuse fuse, const Ops<RS>* __restrict__ gops)
{
  using L = Euler_lds<ND, RS, DEF, P>;
  constexpr int EPB = Shape<ND, RS>::EPB;
  __shared__ typename P::T lt[EPB * L::PER_T];
  __shared__ double ld[EPB * L::PER_D];
  __shared__ __align__(16) float mtab[P::mfma ? 4 * Eops<RS, double>::MF_STRIDE : 4];
  local_euler_tile<ND, RS, DEF, P, FOLD>(m, blockIdx.x, lt, ld, mtab, o, od, coef, dtp, stage, compute_residual,
                                         use_filter, fuse, gops);
}

// ---- Local, Navier-Stokes stage 0 (Spatial.hpp:326-509 with Navier_stokes<true>) -------------------------
template <int ND, int RS, bool DEF, class P>
struct Ns_lds
{
  using S = Shape<ND, RS>;
  static constexpr int NQP = S::NQP, NFQ = S::NFQ, NV = S::NV;
  // arrays of T (padded), gradient phase. FP64 form: the state, one gradient component and the column of the
  // reference-level normals that component j needs, contracted row by row. FP32-MFMA form: the face-referenced
  // products u n_ij of the directions i that contribute to one physical component j, contracted in place.
  static constexpr int NI = DEF ? ND : 1;                         // reference directions per physical component
  static constexpr int U = 0;                                     // [NV][NQP] state (FP64 form)
  static constexpr int G = U + NV * NQP;                          // [NV][NQP] one gradient component (FP64 form)
  static constexpr int SN = G + NV * NQP;                         // [ND][NQP] normals n_ij of one j (FP64 form)
  static constexpr int X = 0;                                     // [NI][NV][NQP] row operands (FP32-MFMA form)
  static constexpr int P1_T = P::mfma ? NI * NV * NQP : SN + (DEF ? ND * NQP : 0);
  // arrays of T (padded), residual phase. FP64 form: one rate array, the viscous rate moving to registers before the
  // convective pass (so that four workgroups of 3D row size 6 fit a compute unit's LDS); FP32-MFMA form: both rates.
  static constexpr int F = 0;                                     // [NV][NQP] viscous, then convective flux
  static constexpr int RT = F + NV * NQP;                         // [NV][NQP] convective rate
  static constexpr int RV = P::mfma ? RT + NV * NQP : RT;         // [NV][NQP] viscous rate
  static constexpr int P2_T = RV + NV * NQP;
  static constexpr int PER_T = P1_T > P2_T ? P1_T : P2_T;
  // arrays of double
  static constexpr int SL = 0;                                    // [2ND][NV][NFQ] LDG face state
  static constexpr int SFN = SL + 2 * ND * NV * NFQ;              // [2ND][ND][NFQ] face normals
  static constexpr int P1_D = SFN + (DEF ? 2 * ND * ND * NFQ : 0);
  static constexpr int SB = 0;                                    // [2ND][NV][NFQ] numerical face fluxes
  static constexpr int PD = SB + 2 * ND * NV * NFQ;               // [ND][NQP] FP32-MFMA, deformed: pressure part p n_iv
  static constexpr int RP = PD + ND * NQP;                        // [ND][NQP] FP32-MFMA, deformed: its derivative, summed
  static constexpr int P2_D = PD + (P::mfma && DEF ? 2 * ND * NQP : 0);
  static constexpr int PER_D = P1_D > P2_D ? P1_D : P2_D;
};

// waves per SIMD that the Navier-Stokes Local's LDS allows, at most 4 (3 for row sizes above 6, which spill at 4):
// the register target follows it
template <int ND, int RS, bool DEF, class P>
constexpr int ns_waves()
{
  using L = Ns_lds<ND, RS, DEF, P>;
  constexpr int lds = Shape<ND, RS>::EPB * (int(sizeof(typename P::T)) * L::PER_T + 8 * L::PER_D);
  constexpr int w = (163840 / lds) * ((form_block<ND, RS, P>() + 63) / 64) / 4, cap = RS > 6 ? 3 : 4;
  return w < 1 ? 1 : (w > cap ? cap : w);
}

// FOLD: the faces' numerical fluxes and LDG averages are computed here from both sides' states (face_flux)
template <int ND, int RS, bool DEF, class P, bool FOLD>
__global__ __launch_bounds__((form_block<ND, RS, P>())) __attribute__((amdgpu_waves_per_eu(ns_waves<ND, RS, DEF, P>())))
void k_local_ns(Dev_mesh m, Eops<RS, double> o,
                const Eops<RS, double>* od, Ns_params ns, double coef, const double* __restrict__ dtp,
                int compute_residual, int use_filter)
{
  using S = Shape<ND, RS>;
  using L = Ns_lds<ND, RS, DEF, P>;
  using T = typename P::T;
  constexpr int NQ = S::NQ, NQP = S::NQP, NFQ = S::NFQ, NV = S::NV, EPB = S::EPB, NT = form_block<ND, RS, P>();
  constexpr int cls = DEF ? HX_DEFORMED : HX_CARTESIAN;
  __shared__ T lt[EPB * L::PER_T];
  __shared__ double ld[EPB * L::PER_D];
  // FP32-MFMA form: the matrix-core operands, read from LDS at each contraction (Contraction_lds)
  __shared__ __align__(16) float mtab[P::mfma ? 4 * Eops<RS, double>::MF_STRIDE : 4];
  // FP64 form, where the LDS keeps the workgroups per compute unit: the elements' face slots, read with the first
  // loads so that the viscous flux extrapolations below need no global read ahead of their writes (3.13 -> 3.09 ms on
  // the 48^3 deformed mesh; the FP32-MFMA form, whose registers are full, lost 1% with it)
  constexpr int LDS_B = EPB * (int(sizeof(T)) * L::PER_T + 8 * L::PER_D), SLOTS_B = 4 * EPB * 2 * ND;
  constexpr bool STAGE = !P::mfma && 163840 / (LDS_B + SLOTS_B) == 163840 / LDS_B;
  __shared__ int fsl[STAGE ? EPB * 2 * ND : 1];
  const int t = threadIdx.x;
  const int blk = blockIdx.x;
  if (STAGE && t < EPB * 2 * ND) {
    const int e2 = blk * EPB + t / (2 * ND);
    fsl[t] = e2 < m.n_elem[cls] ? m.face_slot[cls][e2 * 2 * ND + t % (2 * ND)] : 0;
  }
  const bool pt = t < EPB * NQ;
  const int el = pt ? t / NQ : 0, q = pt ? t % NQ : 0, qp = S::pad(q);
  const int n = m.n_elem[cls];
  const int e = blk * EPB + el;
  const bool valid = pt && e < n;
  T* my = lt + el * L::PER_T;
  if constexpr (P::mfma) lds_mfma_operands<RS, NT>(od, mtab);
  const auto con = [&] {
    if constexpr (P::mfma) return Contraction_lds<RS>(mtab);
    else return Contraction<RS, P, true>(o, od);
  }();

  // the numerical face fluxes are needed only after the gradient; their loads are issued now and held in
  // registers so that their latency hides behind the gradient phase
  constexpr int PF = 2 * ND * NV * NFQ, NSB = FOLD ? 1 : (EPB * PF + NT - 1) / NT;
  double sb[NSB];
  if constexpr (!FOLD) {
#pragma unroll
    for (int k = 0; k < NSB; ++k) {
      const int kk = t + k * NT, el2 = kk / PF, r = kk % PF, f = r / (NV * NFQ), vq = r % (NV * NFQ);
      const int e2 = blk * EPB + el2;
      sb[k] = (kk < EPB * PF && e2 < n) ? state_ptr(m, m.face_slot[cls][e2 * 2 * ND + f], m.alt_rd)[vq] : 0.;
    }
  }
  // folded: one thread per face point computes its flux (held in registers until the convective pass, whose LDS it
  // shares with the gradient phase) and its LDG average (into LDS now)
  constexpr int NFP = 2 * ND * NFQ, FPT = FOLD ? (EPB * NFP + NT - 1) / NT : 1;
  double sbf[FPT][NV];
  if constexpr (FOLD) {
#pragma unroll
    for (int k = 0; k < FPT; ++k) {
      const int kk = t + k * NT, el2 = kk / NFP, r = kk % NFP, face = r / NFQ, p = r % NFQ;
      const int e2 = blk * EPB + el2;
      double avg[NV];
      if (kk < EPB * NFP && e2 < n) {
        face_flux<ND, RS, true>(m, cls, e2, face, p, sbf[k], avg);
      } else {
        for (int v = 0; v < NV; ++v) sbf[k][v] = avg[v] = 0.;
      }
      if (kk < EPB * NFP)
        for (int v = 0; v < NV; ++v) ld[el2 * L::PER_D + L::SL + (face * NV + v) * NFQ + p] = avg[v];
    }
  }

  double u[NV];
  const double* st = m.state[cls] + size_t(e) * NV * NQ;
  for (int v = 0; v < NV; ++v) {
    u[v] = valid ? st[v * NQ + q] : (v >= ND ? 1. : 0.);
    if (!P::mfma && pt) my[L::U + v * NQP + qp] = T(u[v]);
  }
  T bulk = 0, lap = 0;
  if (valid && m.av[cls]) {
    bulk = T(m.av[cls][size_t(e) * 2 * NQ + q]);
    lap = T(m.av[cls][size_t(e) * 2 * NQ + NQ + q]);
  }
  double det = 1.;
  double metd[DEF ? ND : 1][DEF ? ND : 1];
  if constexpr (DEF) {
    const double* mt = m.metrics + size_t(e) * (ND * ND + 1) * NQ;
    for (int i = 0; i < ND; ++i)
      for (int j = 0; j < ND; ++j) {
        metd[i][j] = valid ? mt[(i * ND + j) * NQ + q] : double(i == j);
        if (!P::mfma && pt && j == 0) my[L::SN + i * NQP + qp] = T(metd[i][j]);
      }
    det = valid ? mt[ND * ND * NQ + q] : 1.;
  }
  auto metric = [&](int i, int j) -> double {
    if constexpr (DEF) return metd[i][j];
    else return double(i == j);
  };
  if constexpr (!FOLD) load_faces<ND, RS, NT>(m, cls, blk, 1, ld + L::SL, L::PER_D);
  if constexpr (DEF) {
    for (int k = t; k < EPB * 2 * ND * ND * NFQ; k += NT) {
      const int el2 = k / (2 * ND * ND * NFQ), r = k % (2 * ND * ND * NFQ), f = r / (ND * NFQ), jq = r % (ND * NFQ);
      const int e2 = blk * EPB + el2;
      double val = 0.;
      if (e2 < n) {
        const int ns_ = m.face_nslot[e2 * 2 * ND + f];
        // a face in a Cartesian connection uses the identity normal (Spatial.hpp:331-339,364-367)
        val = ns_ >= 0 ? m.normals[size_t(ns_) * ND * NFQ + jq] : double(jq / NFQ == f / 2);
      }
      ld[el2 * L::PER_D + L::SFN + r] = val;
    }
  }
  __syncthreads();
  // gradient times Jacobian determinant, divided by the nominal size (Spatial.hpp:371-402), one physical
  // component j at a time
  const double inv_h = valid ? 1. / m.nom_size[cls][e] : 1.;
  T grad[NV][ND];
  if constexpr (!P::mfma) {
    for (int j = 0; j < ND; ++j) {
      for (int i = DEF ? 0 : j; i < (DEF ? ND : j + 1); ++i) {
        for_columns<ND, RS, NV, NT, false>(i, [&](bool ok, int el2, int v, int r, int base, int sd) {
          T* b = lt + el2 * L::PER_T;
          const double* db = ld + el2 * L::PER_D;
          T x[RS], y[RS];
          load_row(b + L::U + v * NQP + base, sd, x);
          double b0 = db[L::SL + ((2 * i) * NV + v) * NFQ + r], b1 = db[L::SL + ((2 * i + 1) * NV + v) * NFQ + r];
          if constexpr (DEF) {
            T nk[RS];
            load_row(b + L::SN + i * NQP + base, sd, nk);
#pragma unroll
            for (int k = 0; k < RS; ++k) x[k] *= nk[k];
            b0 *= db[L::SFN + ((2 * i) * ND + j) * NFQ + r];
            b1 *= db[L::SFN + ((2 * i + 1) * ND + j) * NFQ + r];
          }
          con.faces(x, b0, b1, y);
          T* out = b + L::G + v * NQP + base;
#pragma unroll
          for (int k = 0; k < RS; ++k) out[k * sd] = (i == 0 || !DEF) ? y[k] : out[k * sd] + y[k];
        });
        __syncthreads();
      }
      for (int v = 0; v < NV; ++v) grad[v][j] = my[L::G + v * NQP + qp] * T(inv_h);
      // the next component's column of the normals (every row of this one was read before the barrier above)
      if constexpr (DEF)
        if (pt && j + 1 < ND)
          for (int i = 0; i < ND; ++i) my[L::SN + i * NQP + qp] = T(metric(i, j + 1));
      __syncthreads();
    }
  } else {
    for (int j = 0; j < ND; ++j) {
      // u n_ij minus its value at the row's face 0 (face state times face normal), in FP64, then rounded
      if (pt) {
        const double* db = ld + el * L::PER_D;
        for (int ii = 0; ii < L::NI; ++ii) {
          const int i = DEF ? ii : j, r = Row<ND, RS>::fq(q, i);
          const double nf = DEF ? db[L::SFN + ((2 * i) * ND + j) * NFQ + r] : 1.;
          const double nq = metric(i, j);
          for (int v = 0; v < NV; ++v)
            my[L::X + (ii * NV + v) * NQP + qp] = T(u[v] * nq - db[L::SL + ((2 * i) * NV + v) * NFQ + r] * nf);
        }
      }
      __syncthreads();
      // every direction's rows in one pass, each row replaced by its derivative (a row is read and written by the
      // one thread that owns it)
      for (int ii = 0; ii < L::NI; ++ii) {
        const int i = DEF ? ii : j;
        for_columns<ND, RS, NV, NT, true>(i, [&](bool ok, int el2, int v, int r, int base, int sd) {
          T* row = lt + el2 * L::PER_T + L::X + (ii * NV + v) * NQP + base;
          const double* db = ld + el2 * L::PER_D;
          double b0 = db[L::SL + ((2 * i) * NV + v) * NFQ + r], b1 = db[L::SL + ((2 * i + 1) * NV + v) * NFQ + r];
          if constexpr (DEF) {
            b0 *= db[L::SFN + ((2 * i) * ND + j) * NFQ + r];
            b1 *= db[L::SFN + ((2 * i + 1) * ND + j) * NFQ + r];
          }
          T x[RS], y[RS];
          load_row(row, sd, x);
          con.faces(x, 0., b1 - b0, y);
          if (ok)
#pragma unroll
            for (int k = 0; k < RS; ++k) row[k * sd] = y[k];
        });
      }
      __syncthreads();
      // (the next component's operands overwrite only this thread's own point)
      for (int v = 0; v < NV; ++v) {
        T g = 0;
        for (int ii = 0; ii < L::NI; ++ii) g += my[L::X + (ii * NV + v) * NQP + qp];
        grad[v][j] = g * T(inv_h);
      }
    }
  }
  if constexpr (DEF) {
    const T inv_det = T(1. / det);
    for (int v = 0; v < NV; ++v)
      for (int j = 0; j < ND; ++j) grad[v][j] *= inv_det;
  }
  // physical viscous flux (Spatial.hpp:420-435); its product with the normals is taken per direction below
  T fvis[NV][ND];
  {
    T ut[NV];
    for (int v = 0; v < NV; ++v) ut[v] = T(u[v]);
    flux_visc<ND>(ns, ut, bulk, lap, grad, fvis);
  }
  if constexpr (FOLD) {
#pragma unroll
    for (int k = 0; k < FPT; ++k) {
      const int kk = t + k * NT, el2 = kk / NFP, r = kk % NFP, face = r / NFQ, p = r % NFQ;
      if (kk < EPB * NFP)
        for (int v = 0; v < NV; ++v) ld[el2 * L::PER_D + L::SB + (face * NV + v) * NFQ + p] = sbf[k][v];
    }
  } else {
#pragma unroll
    for (int k = 0; k < NSB; ++k)
      if (t + k * NT < EPB * PF) {
        const int kk = t + k * NT;
        ld[(kk / PF) * L::PER_D + L::SB + kk % PF] = sb[k];
      }
  }
  // residual along each direction (Spatial.hpp:444-470). The viscous flux goes first (it is also extrapolated to the
  // LDG blocks of the faces), then the convective flux with its numerical face fluxes. The two rates accumulate in
  // separate arrays, each over the directions in order, so the order of the two passes leaves the result unchanged;
  // it frees the viscous flux's registers before the convective pass.
  for (int i = 0; i < ND; ++i) {
    __syncthreads();
    if (pt)
      for (int v = 0; v < NV; ++v) {
        T fv;
        if constexpr (DEF) {
          fv = 0;
          for (int j = 0; j < ND; ++j) fv += fvis[v][j] * T(metric(i, j));
        } else {
          fv = fvis[v][i];
        }
        my[L::F + v * NQP + qp] = fv;
      }
    __syncthreads();
    for_columns<ND, RS, NV, NT, P::mfma>(i, [&](bool ok, int el2, int v, int r, int base, int sd) {
      T* b = lt + el2 * L::PER_T;
      T x[RS], y[RS], bx[2];
      load_row(b + L::F + v * NQP + base, sd, x);
      con.faces_if(false, x, 0., 0., y, bx);
      const int e2 = blk * EPB + el2;
      if (ok && e2 < n) {
        const int* fs = (STAGE ? fsl + el2 * 2 * ND : m.face_slot[cls] + e2 * 2 * ND) + 2 * i;
        slot_ptr(m, fs[0])[(NV + v) * NFQ + r] = double(bx[0]);
        slot_ptr(m, fs[1])[(NV + v) * NFQ + r] = double(bx[1]);
      }
      if (ok) sub_row(b + L::RV + v * NQP + base, sd, y, i == 0);
    });
  }
  // FP64 form: the viscous rate (filtered) moves to registers, its arrays take the convective rate
  T vrate[P::mfma ? 1 : NV];
  if constexpr (!P::mfma) {
    if (use_filter) filter_columns<ND, RS, NV, NT>(o, lt + L::RV, L::PER_T);
    else __syncthreads();
    for (int v = 0; v < NV; ++v) vrate[v] = my[L::RV + v * NQP + qp];
  }
  const double p64 = P::mfma ? pressure_of<ND>(u) : 0.;
  // the time step scale of the stage update, loaded with the last direction's convective pass so that its latency
  // overlaps that pass
  double tss_q = 0.;
  for (int i = 0; i < ND; ++i) {
    if (i == ND - 1 && valid) tss_q = m.tss[cls][size_t(e) * NQ + q];
    double nrmd[ND];
    for (int j = 0; j < ND; ++j) nrmd[j] = metric(i, j);
    __syncthreads();
    if (pt)
      conv_point<ND, DEF, P>(u, p64, nrmd, ld + el * L::PER_D + L::SB + (2 * i) * NV * NFQ + Row<ND, RS>::fq(q, i), NFQ,
                             my + L::F + qp, ld + el * L::PER_D + L::PD + qp, NQP);
    __syncthreads();
    for_columns<ND, RS, NV, NT, P::mfma>(i, [&](bool ok, int el2, int v, int r, int base, int sd) {
      T* b = lt + el2 * L::PER_T;
      double* db = ld + el2 * L::PER_D;
      T y[RS];
      conv_row<ND, RS, DEF, P>(con, Ops_reload<RS>{od}, v, b + L::F + v * NQP + base, db + L::PD + (v % ND) * NQP + base, sd,
                               db[L::SB + ((2 * i) * NV + v) * NFQ + r], db[L::SB + ((2 * i + 1) * NV + v) * NFQ + r], y);
      if (ok) {
        sub_row(b + L::RT + v * NQP + base, sd, y, i == 0);
        pressure_row<ND, RS, DEF, P>(Ops_reload<RS>{od}, i, v, db + L::PD + (v % ND) * NQP + base, db + L::RP + (v % ND) * NQP + base, sd);
      }
    });
  }
  if (use_filter) {
    filter_columns<ND, RS, (P::mfma ? 2 : 1) * NV, NT>(o, lt + L::RT, L::PER_T);
    if constexpr (P::mfma && DEF) filter_columns<ND, RS, ND, NT>(o, ld + L::RP, L::PER_D);
  } else {
    __syncthreads();
  }
  // stage-0 update (Spatial.hpp:483-503): the convective part goes to the residual cache
  if (valid) {
    double mult = step_coef(coef, dtp) * tss_q * inv_h;
    if constexpr (DEF) mult /= det;
    double* rc = m.rcache[cls] + size_t(e) * S::NRC * NQ;
    double* sw = m.state[cls] + size_t(e) * NV * NQ;
    for (int v = 0; v < NV; ++v) {
      double r = double(my[L::RT + v * NQP + qp]);
      if (has_pressure<ND, DEF>(v) && P::mfma) r -= ld[el * L::PER_D + L::RP + v * NQP + qp];
      rc[v * NQ + q] = r;
      if constexpr (P::mfma) r += double(my[L::RV + v * NQP + qp]);
      else r += vrate[v];
      r *= mult;
      if (compute_residual) rc[v * NQ + q] = r;
      else sw[v * NQ + q] = u[v] + r;
    }
  }
}

// ---- Reconcile_ldg_flux (Spatial.hpp:543-594) ------------------------------------------------------------
template <int ND, int RS>
struct Reconcile_lds
{
  using S = Shape<ND, RS>;
  static constexpr int SL = 0;                                // [2ND][NV][NFQ] viscous flux jumps
  static constexpr int U = SL + 2 * ND * S::NV * S::NFQ;      // [NV][NQP] rate, then the updated state
  static constexpr int PER = U + S::NV * S::NQP;
};

// waves per SIMD that Reconcile_ldg_flux's LDS allows, at most 8: the register target follows it
template <int ND, int RS>
constexpr int reconcile_waves()
{
  constexpr int lds = Shape<ND, RS>::EPB * 8 * Reconcile_lds<ND, RS>::PER;
  constexpr int w = (163840 / lds) * ((Shape<ND, RS>::BLOCK + 63) / 64) / 4;
  return w < 1 ? 1 : (w > 8 ? 8 : w);
}

template <int ND, int RS>
__device__ __forceinline__ void ldg_jump(const Dev_mesh& m, int cls, int e, int face, int p, double (&out)[ND + 2]);

// FOLD: the jumps are formed here from both sides' LDG blocks (ldg_jump) instead of read from the slots
template <int ND, int RS, bool DEF, bool FOLD = false>
__global__ __launch_bounds__((Shape<ND, RS>::BLOCK)) __attribute__((amdgpu_waves_per_eu(reconcile_waves<ND, RS>())))
void k_reconcile_ldg(Dev_mesh m, Eops<RS, double> o, double dt, const double* __restrict__ dtp, int compute_residual,
                     int use_filter)
{
  using S = Shape<ND, RS>;
  using R = Row<ND, RS>;
  using L = Reconcile_lds<ND, RS>;
  constexpr int NQ = S::NQ, NFQ = S::NFQ, NV = S::NV, EPB = S::EPB;
  constexpr int cls = DEF ? HX_DEFORMED : HX_CARTESIAN;
  __shared__ double lds[EPB * L::PER];
  const int t = threadIdx.x, el = t / NQ, q = t % NQ;
  const int blk = blockIdx.x;
  const int n = m.n_elem[cls];
  const int e = blk * EPB + el;
  const bool valid = e < n;
  double* my = lds + el * L::PER;
  double u[NV];
  const double* st = m.state[cls] + size_t(e) * NV * NQ;
  for (int v = 0; v < NV; ++v) u[v] = valid ? st[v * NQ + q] : 0.;
  if constexpr (FOLD) {
    constexpr int NFP = 2 * ND * NFQ;
    for (int k = t; k < EPB * NFP; k += S::BLOCK) {
      const int el2 = k / NFP, r = k % NFP, face = r / NFQ, p = r % NFQ;
      const int e2 = blk * EPB + el2;
      double jv[NV];
      if (e2 < n) ldg_jump<ND, RS>(m, cls, e2, face, p, jv);
      else
        for (int v = 0; v < NV; ++v) jv[v] = 0.;
      for (int v = 0; v < NV; ++v) lds[el2 * L::PER + L::SL + (face * NV + v) * NFQ + p] = jv[v];
    }
  } else {
    load_faces<ND, RS, S::BLOCK>(m, cls, blk, 1, lds + L::SL, L::PER);
  }
  __syncthreads();
  // -boundary_term(jump) (Spatial.hpp:560-570)
  double rate[NV] = {};
  for (int i = 0; i < ND; ++i) {
    const int ki = R::node(q, i), fq = R::fq(q, i);
    const double l0 = o.lift[ki][0], l1 = o.lift[ki][1];
    for (int v = 0; v < NV; ++v)
      rate[v] -= l0 * my[L::SL + ((2 * i) * NV + v) * NFQ + fq] + l1 * my[L::SL + ((2 * i + 1) * NV + v) * NFQ + fq];
  }
  if (use_filter) {
    for (int v = 0; v < NV; ++v) my[L::U + v * S::NQP + S::pad(q)] = rate[v];
    filter_columns<ND, RS, NV, S::BLOCK>(o, lds + L::U, L::PER);
    for (int v = 0; v < NV; ++v) rate[v] = my[L::U + v * S::NQP + S::pad(q)];
  }
  if (valid) {
    double mult = step_coef(dt, dtp) * m.tss[cls][size_t(e) * NQ + q] / m.nom_size[cls][e];
    if constexpr (DEF) mult /= m.metrics[size_t(e) * (ND * ND + 1) * NQ + ND * ND * NQ + q];
    if (compute_residual) {
      double* rc = m.rcache[cls] + size_t(e) * S::NRC * NQ;
      for (int v = 0; v < NV; ++v) rc[v * NQ + q] += rate[v] * mult;
    } else {
      double* sw = m.state[cls] + size_t(e) * NV * NQ;
      for (int v = 0; v < NV; ++v) {
        u[v] += rate[v] * mult;
        sw[v * NQ + q] = u[v];
      }
    }
  }
  // Write_face of the updated state (Spatial.hpp:589)
  __syncthreads();
  for (int v = 0; v < NV; ++v) my[L::U + v * S::NQP + S::pad(q)] = u[v];
  __syncthreads();
  write_faces_from<ND, RS, S::BLOCK>(m, cls, blk, o, lds + L::U, L::PER);
}

// ---- Write_face (Spatial.hpp:60-70) ----------------------------------------------------------------------
// Write_face (Spatial.hpp:29-71) of the state: staged in LDS in the padded layout, then one thread per column of each
// direction in turn extrapolating it to both of that direction's faces (write_faces_from, the operator constants
// taken by value). scale_velocity 1: first divide the momentum by sqrt(2 rho E) in the state itself, the advection
// velocity of Solver::update_art_visc_smoothness (Solver.cpp:468-478); 2: multiply it back (Solver.cpp:570-574).
// Either way the Write_face that follows it in the reference reads the state once.
template <int ND, int RS>
__global__ __launch_bounds__((Shape<ND, RS>::BLOCK)) void k_write_face(Dev_mesh m, Eops<RS, double> o, int cls,
                                                                       int scale_velocity)
{
  using S = Shape<ND, RS>;
  constexpr int NQ = S::NQ, NQP = S::NQP, NV = S::NV, EPB = S::EPB;
  __shared__ double su[EPB * NV * NQP];
  const int t = threadIdx.x, el = t / NQ, q = t % NQ;
  const int e = blockIdx.x * EPB + el;
  const bool valid = e < m.n_elem[cls];
  double* st = m.state[cls] + size_t(e) * NV * NQ;
  double u[NV];
  for (int v = 0; v < NV; ++v) u[v] = valid ? st[v * NQ + q] : 0.;
  if (scale_velocity && valid) {
    const double scale = sqrt(2 * u[ND] * u[ND + 1]);
    for (int d = 0; d < ND; ++d) st[d * NQ + q] = u[d] = scale_velocity == 1 ? u[d] / scale : u[d] * scale;
  }
  for (int v = 0; v < NV; ++v) su[el * NV * NQP + v * NQP + S::pad(q)] = u[v];
  __syncthreads();
  write_faces_from<ND, RS, S::BLOCK>(m, cls, blockIdx.x, o, su, NV * NQP);
}

// ---- Neighbor_reconcile (Spatial.hpp:716-759) ------------------------------------------------------------
// face point f of connection con of class DEF
template <int ND, int RS, bool DEF, int NVAR>
__device__ __forceinline__ void neighbor_reconcile_con(const Dev_mesh& m, int con, int f)
{
  constexpr int NFQ = Shape<ND, RS>::NFQ, NV = ND + 2;  // LDG block offset NV * NFQ; NVAR variables reconciled
  constexpr int cls = DEF ? HX_DEFORMED : HX_CARTESIAN;
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
  for (int v = 0; v < NVAR; ++v) {
    const double a0 = f0[v * NFQ + f], a1 = f1[v * NFQ + p1];
    double avg = 0;
    avg += .5 * sg0 * a0;
    avg += .5 * sg1 * a1;
    f0[v * NFQ + f] = sg0 * avg - a0;
    f1[v * NFQ + p1] = sg1 * avg - a1;
  }
}

template <int ND, int RS, bool DEF, int NVAR = ND + 2>
__global__ __launch_bounds__(256) void k_neighbor_reconcile(Dev_mesh m)
{
  constexpr int NFQ = Shape<ND, RS>::NFQ;
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int con = gid / NFQ, f = gid % NFQ;
  if (con < m.n_con[DEF ? HX_DEFORMED : HX_CARTESIAN]) neighbor_reconcile_con<ND, RS, DEF, NVAR>(m, con, f);
}

// Neighbor_reconcile over the connections that touch a hanging-face mortar (Dev_mesh::mortar_con), the ones whose
// jumps the folded Reconcile_ldg_flux does not form itself
template <int ND, int RS, bool DEF>
__global__ __launch_bounds__(256) void k_neighbor_reconcile_mortar(Dev_mesh m)
{
  constexpr int NFQ = Shape<ND, RS>::NFQ;
  constexpr int cls = DEF ? HX_DEFORMED : HX_CARTESIAN;
  const int gid = blockIdx.x * 256 + threadIdx.x;
  const int i = gid / NFQ, f = gid % NFQ;
  if (i < m.n_mortar_con[cls]) neighbor_reconcile_con<ND, RS, DEF, ND + 2>(m, m.mortar_con[cls][i], f);
}

// Neighbor_reconcile folded into Reconcile_ldg_flux: the reconciled viscous-flux jump of face point p of face `face`
// of element e, this side's value of what Neighbor_reconcile writes (the same operations on the two sides' LDG blocks,
// in the connection's side order), from the partner table (Dev_mesh::partner). A face whose partner is -1 (hanging
// faces) reads the jump a launch before wrote into its own slot.
template <int ND, int RS>
__device__ __forceinline__ void ldg_jump(const Dev_mesh& m, int cls, int e, int face, int p, double (&out)[ND + 2])
{
  constexpr int NFQ = Shape<ND, RS>::NFQ, NV = ND + 2;
  const size_t k = size_t(e) * 2 * ND + face;
  const int* pi = m.partner[cls] + 4 * k;
  const double* own = slot_ptr(m, m.face_slot[cls][k]) + NV * NFQ;
  const int ps = pi[0];
  if (ps < 0) {
    for (int v = 0; v < NV; ++v) out[v] = own[v * NFQ + p];
    return;
  }
  const int info = pi[1];
  const Dir dir(info & 63);
  const bool side1 = (info >> 6) & 1, def = (info >> 7) & 1;
  int fp = p;
  double sg0 = 1., sg1 = 1.;
  if (def) {
    fp = side1 ? face0_index<ND, RS>(dir, p) : face1_index<ND, RS>(dir, p);
    sg0 = 1 - 2 * int(dir.flip_normal(0));
    sg1 = 1 - 2 * int(dir.flip_normal(1));
  }
  const double* par = slot_ptr(m, ps) + NV * NFQ;
  for (int v = 0; v < NV; ++v) {
    const double ao = own[v * NFQ + p], ap = par[v * NFQ + fp];
    const double a0 = side1 ? ap : ao, a1 = side1 ? ao : ap;
    double avg = 0;
    avg += .5 * sg0 * a0;
    avg += .5 * sg1 * a1;
    out[v] = side1 ? sg1 * avg - a1 : sg0 * avg - a0;
  }
}

// ---- Max_dt (Spatial.hpp:784-828): per-point scale; grid-stride loop, wavefront + workgroup min and one atomic
// per workgroup, with the grid capped so that the atomics on the single result word stay few ----
constexpr int max_dt_blocks = 2048;

template <int ND, int RS, bool VISC>
__device__ inline double max_dt_scale(const Dev_mesh& m, const Ops<RS>* __restrict__ gops, const Ns_params& ns, int cls,
                                      size_t e, int q, double max_cfl_c, double max_cfl_d)
{
  constexpr int NQ = Shape<ND, RS>::NQ, NV = ND + 2;
  double u[NV];
  const double* st = m.state[cls] + e * NV * NQ;
  for (int v = 0; v < NV; ++v) u[v] = st[v * NQ + q];
  return max_dt_scale_u<ND, RS, VISC>(m, gops, ns, cls, e, q, u, max_cfl_c, max_cfl_d);
}

template <int ND, int RS, bool VISC>
__global__ __launch_bounds__(256) void k_max_dt(Dev_mesh m, const Ops<RS>* __restrict__ gops, Ns_params ns, int cls,
                                                int is_local, double max_cfl_c, double max_cfl_d, double* result)
{
  constexpr int NQ = Shape<ND, RS>::NQ;
  __shared__ double wmin[256 / 64];
  const size_t n_items = size_t(m.n_elem[cls]) * NQ;
  double inv = DBL_MAX;
  for (size_t gid = size_t(blockIdx.x) * 256 + threadIdx.x; gid < n_items; gid += size_t(gridDim.x) * 256) {
    const size_t e = gid / NQ;
    const int q = int(gid % NQ);
    const double scale = max_dt_scale<ND, RS, VISC>(m, gops, ns, cls, e, q, max_cfl_c, max_cfl_d);
    if (is_local) {
      m.tss[cls][e * NQ + q] = 1. / scale;
    } else {
      m.tss[cls][e * NQ + q] = 1.;
      const double c = 1. / scale;
      // NaN never wins, as with std::min(dt, NaN) in the reference
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

// atomic minimum of a positive double into case_min[k] (the bits of positive doubles order like the doubles)
__device__ inline void flush_case_min(unsigned long long* case_min, int k, double v)
{
  if (k >= 0 && v < DBL_MAX) atomicMin(case_min + k, (unsigned long long)__double_as_longlong(v));
}

// Global-time-step Max_dt of a batch of independent cases: as k_max_dt, with the minimum taken per case (element e
// belongs to case case_id[e]). Each workgroup scans one contiguous chunk of the points, so that a thread meets few
// case changes: it keeps a running minimum and flushes it into its case's slot whenever the case changes, and at the
// end once per wavefront where the whole wavefront shares one case.
template <int ND, int RS, bool VISC>
__global__ __launch_bounds__(256) void k_max_dt_cases(Dev_mesh m, const Ops<RS>* __restrict__ gops, Ns_params ns,
                                                      int cls, double max_cfl_c, double max_cfl_d,
                                                      const int* __restrict__ case_id,
                                                      unsigned long long* __restrict__ case_min)
{
  constexpr int NQ = Shape<ND, RS>::NQ;
  const size_t n_items = size_t(m.n_elem[cls]) * NQ;
  const size_t chunk = (n_items + gridDim.x - 1) / gridDim.x;
  const size_t lo = size_t(blockIdx.x) * chunk;
  const size_t hi = lo + chunk < n_items ? lo + chunk : n_items;
  int cur = -1;
  double inv = DBL_MAX;
  for (size_t gid = lo + threadIdx.x; gid < hi; gid += 256) {
    const size_t e = gid / NQ;
    const int q = int(gid % NQ);
    const double c = 1. / max_dt_scale<ND, RS, VISC>(m, gops, ns, cls, e, q, max_cfl_c, max_cfl_d);
    m.tss[cls][e * NQ + q] = 1.;
    const int k = case_id[e];
    if (k != cur) {
      flush_case_min(case_min, cur, inv);
      cur = k;
      inv = DBL_MAX;
    }
    inv = (c < inv) ? c : inv;
  }
  const int c0 = __shfl(cur, 0, 64);
  if (__all(cur == c0)) {
    for (int off = 32; off > 0; off /= 2) {
      const double o = __shfl_xor(inv, off, 64);
      inv = (o < inv) ? o : inv;
    }
    if ((threadIdx.x & 63) == 0) flush_case_min(case_min, cur, inv);
  } else {
    flush_case_min(case_min, cur, inv);
  }
}

// ---- Prolong_refined / Restrict_refined (Spatial.hpp:153-285) --------------------------------------------
// The tangential operators of one fine face at one face point (a, b): for each tangential dimension d in order, the
// matrix row of the point's coordinate along d (the row loops of Spatial.hpp:183-199 and 259-275), or for a
// stretched dimension the factor fac, multiplied (prolongation) or divided (restriction). Per output the products
// are summed in the reference's order: dimension 0 first, ascending k.
template <int RS, bool DIVIDE>
struct Face_rows
{
  double r0[2][RS];  // row a of the dimension-0 matrix of either half
  double r1[2][RS];  // row b of the dimension-1 matrix of either half
  bool str[2];
  double fac;
  __device__ double scaled(double x) const { return DIVIDE ? x / fac : x * fac; }
  // (the halves are selected element by element, so that no register array is indexed at run time)
  // dimension-0 result at (a, k) for every k, from face data src (rs x rs, dimension 0 slowest)
  __device__ void dim0(const double* src, int a, bool h0, double (&row)[RS]) const
  {
#pragma unroll
    for (int k = 0; k < RS; ++k) {
      if (str[0]) {
        row[k] = scaled(src[a * RS + k]);
      } else {
        double acc = 0.;
#pragma unroll
        for (int j = 0; j < RS; ++j) acc += (h0 ? r0[1][j] : r0[0][j]) * src[j * RS + k];
        row[k] = acc;
      }
    }
  }
  // the 3D face value at b from the dimension-0 result of the face's half along dimension 0 (row_h0)
  __device__ double dim1(const double (&row_0)[RS], const double (&row_1)[RS], bool h0, int b, bool h1) const
  {
    double x[RS];
#pragma unroll
    for (int k = 0; k < RS; ++k) x[k] = h0 ? row_1[k] : row_0[k];
    if (str[1]) {
      double xb = x[0];
#pragma unroll
      for (int k = 1; k < RS; ++k) xb = b == k ? x[k] : xb;
      return scaled(xb);
    }
    double acc = 0.;
#pragma unroll
    for (int k = 0; k < RS; ++k) acc += (h1 ? r1[1][k] : r1[0][k]) * x[k];
    return acc;
  }
  // the 2D face value at a (one tangential dimension)
  __device__ double line(const double* src, int a, bool h0) const
  {
    if (str[0]) return scaled(src[a]);
    double acc = 0.;
#pragma unroll
    for (int j = 0; j < RS; ++j) acc += (h0 ? r0[1][j] : r0[0][j]) * src[j];
    return acc;
  }
};

// fine faces of refined face i_ref (fewer along stretched dimensions), the stretch flags, and the half (0/1) that
// tangential dimension d of fine face i_face takes (Spatial.hpp:175-181)
template <int ND>
__device__ inline int fine_faces(const Dev_mesh& m, int i_ref, bool (&str)[2])
{
  str[0] = ND > 1 && m.ref_stretch[2 * i_ref];
  str[1] = ND > 2 && m.ref_stretch[2 * i_ref + 1];
  int nf = 1 << (ND - 1);
  for (int d = 0; d < ND - 1; ++d) nf /= 1 + str[d];
  return nf;
}
template <int ND>
__device__ inline int half_of(int i_face, int d, const bool (&str)[2])
{
  if constexpr (ND < 2) {
    return 0;
  } else {
    const int face_stride = str[ND - 2] ? 1 : (1 << (ND - 2 - d));
    return (i_face / face_stride) % 2;
  }
}

// the matrix rows of face point p (coordinates a, b) of both halves, from mats[2][rs][rs]
template <int ND, int RS, bool DIVIDE>
__device__ inline void face_rows(Face_rows<RS, DIVIDE>& fr, const double (*mats)[RS][RS], int a, int b, int scale)
{
  fr.fac = 1. + scale;
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int k = 0; k < RS; ++k) {
      fr.r0[h][k] = ND > 1 ? mats[h][a][k] : 0.;
      fr.r1[h][k] = ND > 2 ? mats[h][b][k] : 0.;
    }
}

// one thread per (refined face, variable, face point): the coarse face is staged in LDS and every fine face's value
// at the point computed from it. In 3D in two phases: each thread first forms the dimension-0 sums at its point
// (a, k) for both halves into LDS (Face_rows::dim0's sums, shared by the fine faces of the same half, so each is
// formed once instead of once per b), then after a barrier its own point's dimension-1 sum for every fine face
// (Face_rows::dim1's). The operator matrices are staged in LDS once per workgroup. The products and their order
// are those of Face_rows.
template <int ND, int RS>
__global__ __launch_bounds__(256) void k_prolong(Dev_mesh m, const Ops<RS>* __restrict__ ops, int n_var, int scale,
                                                 int offset)
{
  constexpr int NFQ = ipow(RS, ND - 1);
  constexpr int G = 256 / NFQ;
  __shared__ double stage[G][NFQ];
  const int g = threadIdx.x / NFQ, p = threadIdx.x % NFQ;
  const size_t gi = size_t(blockIdx.x) * G + g;
  const bool active = g < G && gi < size_t(m.n_ref) * n_var;
  const int i_ref = active ? int(gi / n_var) : 0, v = active ? int(gi % n_var) : 0;
  // block 0 of the slots in the buffer being written (the prolongation follows the Local), or the LDG block
  auto at = [&](int slot) { return offset ? slot_ptr(m, slot) + (ND + 2 + v) * NFQ : state_ptr(m, slot, m.alt_wr) + v * NFQ; };
  if (active) stage[g][p] = at(m.ref_slot[5 * i_ref])[p];
  if constexpr (ND == 3) {
    __shared__ double mat[2][RS][RS];
    __shared__ double rows[G][2][NFQ];
    for (int k = threadIdx.x; k < 2 * RS * RS; k += 256) (&mat[0][0][0])[k] = (&ops->prolong[0][0][0])[k];
    bool str[2] = {false, false};
    const int nf = active ? fine_faces<ND>(m, i_ref, str) : 0;
    const double fac = 1. + scale;
    const int a = p / RS, c = p % RS;  // phase 1: point (a, k = c); phase 2: point (a, b = c)
    __syncthreads();
    if (active) {
      // row h: the dimension-0 half 0, and half 1 unless dimension 0 is stretched (then both are the scaled row)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        double val;
        if (str[0]) {
          val = stage[g][a * RS + c] * fac;
        } else {
          const int half = h;
          val = 0.;
#pragma unroll
          for (int j = 0; j < RS; ++j) val += mat[half][a][j] * stage[g][j * RS + c];
        }
        rows[g][h][p] = val;
      }
    }
    __syncthreads();
    if (!active) return;
    for (int f = 0; f < nf; ++f) {
      const bool h0 = half_of<ND>(f, 0, str), h1 = half_of<ND>(f, 1, str);
      const double* x = rows[g][h0 && !str[0] ? 1 : 0] + a * RS;
      double r;
      if (str[1]) {
        r = x[c] * fac;
      } else {
        r = 0.;
#pragma unroll
        for (int k = 0; k < RS; ++k) r += mat[h1 ? 1 : 0][c][k] * x[k];
      }
      at(m.ref_slot[5 * i_ref + 1 + f])[p] = r;
    }
  } else {
    __syncthreads();
    if (!active) return;
    Face_rows<RS, false> fr;
    const int nf = fine_faces<ND>(m, i_ref, fr.str);
    const int a = ND > 2 ? p / RS : p, b = ND > 2 ? p % RS : 0;
    face_rows<ND, RS, false>(fr, ops->prolong, a, b, scale);
    if constexpr (ND == 1) {
      at(m.ref_slot[5 * i_ref + 1])[p] = stage[g][p];
    } else {
      for (int f = 0; f < nf; ++f) at(m.ref_slot[5 * i_ref + 1 + f])[p] = fr.line(stage[g], a, half_of<ND>(f, 0, fr.str));
    }
  }
}

// one thread per (refined face, variable, face point): the fine faces are staged in LDS, restricted in place
// (Restrict_refined leaves the restricted data in the fine faces) and summed into the coarse face in fine-face order.
// In 3D in two phases as k_prolong: each thread forms the dimension-0 sums at its point (a, k) of every fine face
// into LDS, then its own point's dimension-1 sums.
template <int ND, int RS>
__global__ __launch_bounds__(256) void k_restrict(Dev_mesh m, const Ops<RS>* __restrict__ ops, int n_var, int scale,
                                                  int offset)
{
  constexpr int NFQ = ipow(RS, ND - 1);
  constexpr int N_FACE = 1 << (ND - 1);
  constexpr int G = 256 / NFQ;
  __shared__ double stage[G][N_FACE][NFQ];
  const int g = threadIdx.x / NFQ, p = threadIdx.x % NFQ;
  const size_t gi = size_t(blockIdx.x) * G + g;
  const bool active = g < G && gi < size_t(m.n_ref) * n_var;
  const int i_ref = active ? int(gi / n_var) : 0, v = active ? int(gi % n_var) : 0;
  Face_rows<RS, true> fr;
  fr.str[0] = fr.str[1] = false;
  const int nf = active ? fine_faces<ND>(m, i_ref, fr.str) : 0;
  // block 0 of the slots in the buffer being read (the restriction precedes the Local), or the LDG block
  auto at = [&](int slot) { return offset ? slot_ptr(m, slot) + (ND + 2 + v) * NFQ : state_ptr(m, slot, m.alt_rd) + v * NFQ; };
  for (int f = 0; f < nf; ++f) stage[g][f][p] = at(m.ref_slot[5 * i_ref + 1 + f])[p];
  if constexpr (ND == 3) {
    __shared__ double mat[2][RS][RS];
    __shared__ double rows[G][N_FACE][NFQ];
    for (int k = threadIdx.x; k < 2 * RS * RS; k += 256) (&mat[0][0][0])[k] = (&ops->restrict_[0][0][0])[k];
    const double fac = 1. + scale;
    const int a = p / RS, c = p % RS;  // phase 1: point (a, k = c); phase 2: point (a, b = c)
    __syncthreads();
    for (int f = 0; f < nf; ++f) {
      double val;
      if (fr.str[0]) {
        val = stage[g][f][a * RS + c] / fac;
      } else {
        const int h0 = half_of<ND>(f, 0, fr.str);
        val = 0.;
#pragma unroll
        for (int j = 0; j < RS; ++j) val += mat[h0][a][j] * stage[g][f][j * RS + c];
      }
      rows[g][f][p] = val;
    }
    __syncthreads();
    if (!active) return;
    double acc = 0.;
    for (int f = 0; f < nf; ++f) {
      const double* x = rows[g][f] + a * RS;
      double r;
      if (fr.str[1]) {
        r = x[c] / fac;
      } else {
        const int h1 = half_of<ND>(f, 1, fr.str);
        r = 0.;
#pragma unroll
        for (int k = 0; k < RS; ++k) r += mat[h1][c][k] * x[k];
      }
      at(m.ref_slot[5 * i_ref + 1 + f])[p] = r;
      acc += r;
    }
    at(m.ref_slot[5 * i_ref])[p] = acc;
  } else {
    __syncthreads();
    if (!active) return;
    const int a = ND > 2 ? p / RS : p, b = ND > 2 ? p % RS : 0;
    face_rows<ND, RS, true>(fr, ops->restrict_, a, b, scale);
    double acc = 0.;
    for (int f = 0; f < nf; ++f) {
      double r;
      if constexpr (ND == 1) r = stage[g][f][p];
      else r = fr.line(stage[g][f], a, half_of<ND>(f, 0, fr.str));
      at(m.ref_slot[5 * i_ref + 1 + f])[p] = r;
      acc += r;
    }
    at(m.ref_slot[5 * i_ref])[p] = acc;
  }
}

// Neighbor on the connections of a refined face's mortars fused with the face's Restrict_refined, for the Euler stages
// of the update path in 3D (scale 1, state blocks): one workgroup per refined face. Phase 1: each (mortar, connection
// point) item forms the connection's numerical flux (con_flux, as Neighbor does) and stores the fine face's side in its
// slot, the mortar's side in LDS instead of the mortar slot; phases 2 and 3: k_restrict's two phases from LDS, storing
// the restricted mortar values and their sum. The values are those of k_neighbor_mortar followed by k_restrict; the
// mortar slots' unrestricted fluxes, which the restriction overwrites, are never written.
template <int ND, int RS>
__global__ __launch_bounds__(256) void k_mortar_restrict(Dev_mesh m, const Ops<RS>* __restrict__ ops)
{
  constexpr int NFQ = ipow(RS, ND - 1), NV = ND + 2, N_FACE = 1 << (ND - 1);
  __shared__ double stage[N_FACE][NV][NFQ];
  __shared__ double rows[N_FACE][NV][NFQ];
  __shared__ double mat[2][RS][RS];
  const int i_ref = blockIdx.x, t = threadIdx.x;
  for (int k = t; k < 2 * RS * RS; k += 256) (&mat[0][0][0])[k] = (&ops->restrict_[0][0][0])[k];
  bool str[2] = {false, false};
  const int nf = fine_faces<ND>(m, i_ref, str);
  for (int it = t; it < nf * NFQ; it += 256) {
    const int f = it / NFQ, q = it % NFQ;
    const int code = m.ref_con[4 * i_ref + f];
    if (code < 0) continue;  // (the host launches this only where every such mortar is in a connection)
    const int con = code & 0x3fffffff;
    const int mortar = m.ref_slot[5 * i_ref + 1 + f];
    auto keep = [&](const Con_flux<ND, RS>& r) {
      if (r.s0 == mortar) {
        for (int v = 0; v < NV; ++v) {
          stage[f][v][q] = r.sg0 * r.fl[v];
          r.f1[v * NFQ + r.p1] = r.sg1 * r.fl[v];
        }
      } else {
        for (int v = 0; v < NV; ++v) {
          r.f0[v * NFQ + q] = r.sg0 * r.fl[v];
          stage[f][v][r.p1] = r.sg1 * r.fl[v];
        }
      }
    };
    if (code >> 30) keep(con_flux<ND, RS, true, false>(m, con, q));
    else keep(con_flux<ND, RS, false, false>(m, con, q));
  }
  constexpr double fac = 2.;  // 1 + scale
  __syncthreads();
  for (int it = t; it < nf * NV * NFQ; it += 256) {
    const int f = it / (NV * NFQ), v = (it / NFQ) % NV, p = it % NFQ;
    const int a = p / RS, c = p % RS;
    double val;
    if (str[0]) {
      val = stage[f][v][a * RS + c] / fac;
    } else {
      const int h0 = half_of<ND>(f, 0, str);
      val = 0.;
#pragma unroll
      for (int j = 0; j < RS; ++j) val += mat[h0][a][j] * stage[f][v][j * RS + c];
    }
    rows[f][v][p] = val;
  }
  __syncthreads();
  for (int it = t; it < NV * NFQ; it += 256) {
    const int v = it / NFQ, p = it % NFQ;
    const int a = p / RS, c = p % RS;
    double acc = 0.;
    for (int f = 0; f < nf; ++f) {
      const double* x = rows[f][v] + a * RS;
      double r;
      if (str[1]) {
        r = x[c] / fac;
      } else {
        const int h1 = half_of<ND>(f, 1, str);
        r = 0.;
#pragma unroll
        for (int k = 0; k < RS; ++k) r += mat[h1][c][k] * x[k];
      }
      state_ptr(m, m.ref_slot[5 * i_ref + 1 + f], m.alt_rd)[v * NFQ + p] = r;
      acc += r;
    }
    state_ptr(m, m.ref_slot[5 * i_ref], m.alt_rd)[v * NFQ + p] = acc;
  }
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

// one thread per (boundary connection, face point) item `gid`
template <int ND, int RS>
__device__ __forceinline__ void state_bcs_at(const Dev_mesh& m, int gid)
{
  constexpr int NFQ = ipow(RS, ND - 1), NV = ND + 2;
  const int ib = gid / NFQ, f = gid % NFQ;
  if (ib >= m.n_bc_con) return;
  const int con = m.bc_con[ib];
  const hx_bc bc = m.bcs[m.bc_serial[ib]];
  const int s_in = m.con_slot[HX_DEFORMED][2 * con], s_gh = m.con_slot[HX_DEFORMED][2 * con + 1];
  // the states (block 0) in the buffer being read
  const double* in_f = state_ptr(m, s_in, m.alt_rd);
  double* gh_f = state_ptr(m, s_gh, m.alt_rd);
  switch (bc.kind) {
    case HX_BC_FREESTREAM:
      for (int v = 0; v < NV; ++v) gh_f[v * NFQ + f] = bc.state[v];
      break;
    case HX_BC_COPY:
    case HX_BC_OUTFLOW:
      for (int v = 0; v < NV; ++v) gh_f[v * NFQ + f] = in_f[v * NFQ + f];
      for (int v = 0; v < NV; ++v) slot_ptr(m, s_gh)[(NV + v) * NFQ + f] = slot_ptr(m, s_in)[(NV + v) * NFQ + f];
      break;
    case HX_BC_NONPENETRATION:
      for (int v = 0; v < NV; ++v) gh_f[v * NFQ + f] = in_f[v * NFQ + f];
      reflect_normal<ND, RS>(gh_f, m.normals + size_t(m.con_nslot[2 * con]) * ND * NFQ, f);
      break;
    case HX_BC_RIEMANN_INVARIANTS:
    case HX_BC_PRESSURE_OUTFLOW:
    case HX_BC_NO_SLIP: {
      const double* nslot = m.normals + size_t(m.con_nslot[2 * con]) * ND * NFQ;
      double* cache = m.bc_cache + size_t(ib) * NV * NFQ;
      const Dir dir(m.con_dir[HX_DEFORMED][con]);
      double in[NV], nrm[ND], gh[NV];
      for (int v = 0; v < NV; ++v) in[v] = in_f[v * NFQ + f];
      for (int d = 0; d < ND; ++d) nrm[d] = nslot[d * NFQ + f];
      if (bc.kind == HX_BC_RIEMANN_INVARIANTS) {
        // Boundary_condition.cpp:97-141: incoming characteristics from the freestream, outgoing from inside, then
        // limited to stay thermodynamically admissible; the inside state primes the cache
        double fs[NV];
        for (int v = 0; v < NV; ++v) fs[v] = bc.state[v];
        apply_char<ND>(in, nrm, 1 - 2 * int(dir.sign0), in, fs, gh);
        gh[ND] = fmax(gh[ND], in[ND] / 2);
        double ke = 0., ike = 0.;
        for (int d = 0; d < ND; ++d) {
          ke += gh[d] * gh[d];
          ike += in[d] * in[d];
        }
        ke = .5 * ke / gh[ND];
        ike = .5 * ike / in[ND];
        gh[ND + 1] = fmax(ke + fmax(gh[ND + 1] - ke, (in[ND + 1] - ike) / 2), 0.);
        for (int v = 0; v < NV; ++v) cache[v * NFQ + f] = in[v];
      } else if (bc.kind == HX_BC_PRESSURE_OUTFLOW) {
        // Boundary_condition.cpp:184-214: the specified pressure where the outflow is subsonic
        double mn = 0., nsq = 0., msq = 0.;
        for (int d = 0; d < ND; ++d) {
          mn += in[d] * nrm[d];
          nsq += nrm[d] * nrm[d];
          msq += in[d] * in[d];
        }
        const double nrml_veloc = mn / in[ND] / sqrt(nsq);
        const double kin = .5 * msq / in[ND];
        const double pres = fmax(.4 * (in[ND + 1] - kin), 0.);
        const double sound = sqrt(1.4 * pres / in[ND]);
        for (int v = 0; v < NV; ++v) gh[v] = in[v];
        if (nrml_veloc * (2 * int(dir.sign0) - 1) < sound) gh[ND + 1] = bc.pressure / .4 + kin;
      } else {
        // Boundary_condition.cpp:381-400: velocity reversed, energy from the thermal condition; the average of the
        // ghost and inside states primes the cache
        for (int d = 0; d < ND; ++d) gh[d] = -in[d];
        gh[ND] = in[ND];
        const double e = ghost_energy<NV>(bc, in);
        gh[ND + 1] = e * e / in[ND + 1];
        for (int v = 0; v < NV; ++v) cache[v * NFQ + f] = (gh[v] + in[v]) / 2;
      }
      for (int v = 0; v < NV; ++v) gh_f[v * NFQ + f] = gh[v];
      break;
    }
    default:
      break;
  }
}

template <int ND, int RS>
__global__ __launch_bounds__(256) void k_state_bcs(Dev_mesh m)
{
  state_bcs_at<ND, RS>(m, blockIdx.x * 256 + threadIdx.x);
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
    case HX_BC_PRESSURE_OUTFLOW:  // Boundary_condition.cpp:216-223
      for (int v = 0; v < NV; ++v) gh_f[(NV + v) * NFQ + f] = -in_f[(NV + v) * NFQ + f];
      break;
    case HX_BC_RIEMANN_INVARIANTS:
    case HX_BC_NO_SLIP: {
      const double* nslot = m.normals + size_t(m.con_nslot[2 * con]) * ND * NFQ;
      const double* cache = m.bc_cache + size_t(ib) * NV * NFQ;
      const Dir dir(m.con_dir[HX_DEFORMED][con]);
      double fl[NV], cs[NV], nrm[ND];
      for (int v = 0; v < NV; ++v) {
        fl[v] = in_f[(NV + v) * NFQ + f];
        cs[v] = cache[v * NFQ + f];
      }
      for (int d = 0; d < ND; ++d) nrm[d] = nslot[d * NFQ + f];
      if (bc.kind == HX_BC_RIEMANN_INVARIANTS) {
        // Boundary_condition.cpp:143-182: the viscous flux of the outgoing characteristics, zero for the incoming
        double zero[NV] = {}, gh[NV];
        apply_char<ND>(cs, nrm, 1 - 2 * int(dir.sign0), zero, fl, gh);
        for (int v = 0; v < NV; ++v) gh_f[(NV + v) * NFQ + f] = gh[v];
      } else {
        // Boundary_condition.cpp:402-427: momentum flux kept, mass flux reversed, energy flux toward the thermal
        // condition's heat flux
        for (int d = 0; d < ND; ++d) gh_f[(NV + d) * NFQ + f] = fl[d];
        gh_f[(NV + ND) * NFQ + f] = -fl[ND];
        double nsq = 0.;
        for (int d = 0; d < ND; ++d) nsq += nrm[d] * nrm[d];
        const double normal = sqrt(nsq);
        const int flux_sign = 2 * int(dir.sign0) - 1;
        const double qe = fl[ND + 1];
        gh_f[(NV + ND + 1) * NFQ + f] =
            bc.coercion * (normal * flux_sign * ghost_heat_flux<NV>(bc, cs, qe * flux_sign / normal) - qe) + qe;
      }
      break;
    }
    default:
      break;
  }
}

}  // namespace hx

#include "hx_av_kernels.hpp"
#include "hx_metrics_kernels.hpp"

namespace hx {

// ---- launchers -------------------------------------------------------------------------------------------
inline unsigned blocks_for(size_t n_items, int per_block) { return unsigned((n_items + per_block - 1) / per_block); }

template <int ND, int RS>
struct Launchers
{
  using S = Shape<ND, RS>;
  static const Ops<RS>* ops(const Launch& l) { return static_cast<const Ops<RS>*>(l.ops); }
  static const Ops<RS>& hops(const Launch& l) { return *static_cast<const Ops<RS>*>(l.hops); }

  static hipError_t write_face(const Launch& l, int cls, int scale_velocity)
  {
    const int n = l.m.n_elem[cls];
    if (n == 0) return hipSuccess;
    const Eops<RS, double> o = make_eops<RS, double>(hops(l));
    hipLaunchKernelGGL((k_write_face<ND, RS>), dim3(blocks_for(n, S::EPB)), dim3(S::BLOCK), 0, l.stream, l.m, o, cls,
                       scale_velocity);
    return hipGetLastError();
  }
  static hipError_t neighbor(const Launch& l, int cls, bool visc)
  {
    if (l.fold) {
      // face-flux folding: only the connections touching a hanging-face mortar
      const int n = l.m.n_mortar_con[cls];
      if (n == 0) return hipSuccess;
      dim3 g(blocks_for(size_t(n) * S::NFQ, 256)), b(256);
      if (cls == HX_DEFORMED) {
        if (visc) hipLaunchKernelGGL((k_neighbor_mortar<ND, RS, true, true>), g, b, 0, l.stream, l.m);
        else hipLaunchKernelGGL((k_neighbor_mortar<ND, RS, true, false>), g, b, 0, l.stream, l.m);
      } else {
        if (visc) hipLaunchKernelGGL((k_neighbor_mortar<ND, RS, false, true>), g, b, 0, l.stream, l.m);
        else hipLaunchKernelGGL((k_neighbor_mortar<ND, RS, false, false>), g, b, 0, l.stream, l.m);
      }
      return hipGetLastError();
    }
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
  template <class P>
  static hipError_t local_form(const Launch& l, int cls, bool visc, double coef, int stage, int cr, int uf)
  {
    const int n = l.m.n_elem[cls];
    if (n == 0) return hipSuccess;
    dim3 g(blocks_for(n, S::EPB)), b(form_block<ND, RS, P>());
    const Eops<RS, double> o = make_eops<RS, double>(hops(l));
    const Eops<RS, double>* od = static_cast<const Eops<RS, double>*>(l.eops);
    if (l.fold) return local_fold<P, true>(l, cls, visc, coef, stage, cr, uf, g, b, o, od);
    return local_fold<P, false>(l, cls, visc, coef, stage, cr, uf, g, b, o, od);
  }
  template <class P, bool FOLD>
  static hipError_t local_fold(const Launch& l, int cls, bool visc, double coef, int stage, int cr, int uf, dim3 g, dim3 b,
                               const Eops<RS, double>& o, const Eops<RS, double>* od)
  {
    if (visc) {
      if (cls == HX_DEFORMED)
        hipLaunchKernelGGL((k_local_ns<ND, RS, true, P, FOLD>), g, b, 0, l.stream, l.m, o, od, l.ns, coef, l.dtp, cr, uf);
      else hipLaunchKernelGGL((k_local_ns<ND, RS, false, P, FOLD>), g, b, 0, l.stream, l.m, o, od, l.ns, coef, l.dtp, cr, uf);
    } else {
      const dim3 be(euler_block<ND, RS>());
      Dt_fuse f = l.fuse;  // the deformed class's workgroups store after the Cartesian class's
      if (f.block_min && cls == HX_DEFORMED) f.block_offset += blocks_for(l.m.n_elem[HX_CARTESIAN], S::EPB);
      if (cls == HX_DEFORMED)
        hipLaunchKernelGGL((k_local_euler<ND, RS, true, euler_form<P, true>, FOLD>), g, be, 0, l.stream, l.m, o, od, coef,
                           l.dtp, stage, cr, uf, f, ops(l));
      else
        hipLaunchKernelGGL((k_local_euler<ND, RS, false, P, FOLD>), g, be, 0, l.stream, l.m, o, od, coef, l.dtp, stage, cr,
                           uf, f, ops(l));
    }
    return hipGetLastError();
  }
  static hipError_t local(const Launch& l, int cls, bool visc, double coef, int stage, int cr, int uf)
  {
    if (l.precision == HX_FP32_MFMA) return local_form<Fp32_mfma_form>(l, cls, visc, coef, stage, cr, uf);
    return local_form<Fp64_form>(l, cls, visc, coef, stage, cr, uf);
  }
  static hipError_t neighbor_reconcile(const Launch& l, int cls)
  {
    if (l.fold) {
      // folded into Reconcile_ldg_flux: only the connections touching a hanging-face mortar
      const int n = l.m.n_mortar_con[cls];
      if (n == 0) return hipSuccess;
      dim3 g(blocks_for(size_t(n) * S::NFQ, 256)), b(256);
      if (cls == HX_DEFORMED) hipLaunchKernelGGL((k_neighbor_reconcile_mortar<ND, RS, true>), g, b, 0, l.stream, l.m);
      else hipLaunchKernelGGL((k_neighbor_reconcile_mortar<ND, RS, false>), g, b, 0, l.stream, l.m);
      return hipGetLastError();
    }
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
    const Eops<RS, double> o = make_eops<RS, double>(hops(l));
    if (l.fold) {
      if (cls == HX_DEFORMED)
        hipLaunchKernelGGL((k_reconcile_ldg<ND, RS, true, true>), g, b, 0, l.stream, l.m, o, dt, l.dtp, cr, uf);
      else
        hipLaunchKernelGGL((k_reconcile_ldg<ND, RS, false, true>), g, b, 0, l.stream, l.m, o, dt, l.dtp, cr, uf);
    } else {
      if (cls == HX_DEFORMED) hipLaunchKernelGGL((k_reconcile_ldg<ND, RS, true>), g, b, 0, l.stream, l.m, o, dt, l.dtp, cr, uf);
      else hipLaunchKernelGGL((k_reconcile_ldg<ND, RS, false>), g, b, 0, l.stream, l.m, o, dt, l.dtp, cr, uf);
    }
    return hipGetLastError();
  }
  static hipError_t max_dt(const Launch& l, int cls, bool visc, int is_local, double cc, double cd)
  {
    const int n = l.m.n_elem[cls];
    if (n == 0) return hipSuccess;
    dim3 g(std::min(blocks_for(size_t(n) * S::NQ, 256), unsigned(max_dt_blocks))), b(256);
    if (visc) hipLaunchKernelGGL((k_max_dt<ND, RS, true>), g, b, 0, l.stream, l.m, ops(l), l.ns, cls, is_local, cc, cd, l.scalar);
    else hipLaunchKernelGGL((k_max_dt<ND, RS, false>), g, b, 0, l.stream, l.m, ops(l), l.ns, cls, is_local, cc, cd, l.scalar);
    return hipGetLastError();
  }
  static hipError_t max_dt_cases(const Launch& l, int cls, bool visc, double cc, double cd, const int* case_id,
                                 unsigned long long* case_min)
  {
    const int n = l.m.n_elem[cls];
    if (n == 0) return hipSuccess;
    dim3 g(std::min(blocks_for(size_t(n) * S::NQ, 256), unsigned(max_dt_blocks))), b(256);
    if (visc)
      hipLaunchKernelGGL((k_max_dt_cases<ND, RS, true>), g, b, 0, l.stream, l.m, ops(l), l.ns, cls, cc, cd, case_id, case_min);
    else
      hipLaunchKernelGGL((k_max_dt_cases<ND, RS, false>), g, b, 0, l.stream, l.m, ops(l), l.ns, cls, cc, cd, case_id, case_min);
    return hipGetLastError();
  }
  static hipError_t prolong(const Launch& l, int n_var, int scale, int offset)
  {
    if (l.m.n_ref == 0) return hipSuccess;
    constexpr int G = 256 / S::NFQ;
    hipLaunchKernelGGL((k_prolong<ND, RS>), dim3(blocks_for(size_t(l.m.n_ref) * n_var, G)), dim3(256), 0, l.stream, l.m,
                       ops(l), n_var, scale, offset);
    return hipGetLastError();
  }
  static hipError_t restrict_(const Launch& l, int n_var, int scale, int offset)
  {
    if (l.m.n_ref == 0) return hipSuccess;
    constexpr int G = 256 / S::NFQ;
    hipLaunchKernelGGL((k_restrict<ND, RS>), dim3(blocks_for(size_t(l.m.n_ref) * n_var, G)), dim3(256), 0, l.stream,
                       l.m, ops(l), n_var, scale, offset);
    return hipGetLastError();
  }
  static hipError_t mortar_restrict(const Launch& l)
  {
    if constexpr (ND != 3) {
      return hipErrorNotSupported;
    } else {
      if (l.m.n_ref == 0) return hipSuccess;
      hipLaunchKernelGGL((k_mortar_restrict<ND, RS>), dim3(l.m.n_ref), dim3(256), 0, l.stream, l.m, ops(l));
      return hipGetLastError();
    }
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
  // artificial-viscosity generator
  static Av_consts<RS> av_consts(const Launch& l, int av_rs)
  {
    const Ops<RS>& o = hops(l);
    Av_consts<RS> c;
    for (int i = 0; i < RS; ++i) {
      c.node[i] = o.node[i];
      c.adv_node[i] = 2 * o.node[i] - 1.;
      c.weight[i] = o.weight[i];
      c.orth[i] = (av_rs >= 1 && av_rs <= RS) ? o.orth[av_rs - 1][i] : 0.;
    }
    return c;
  }
  static hipError_t av_write_face(const Launch& l, int cls, int kind)
  {
    const int n = l.m.n_elem[cls];
    if (n == 0) return hipSuccess;
    const Eops<RS, double> o = make_eops<RS, double>(hops(l));
    dim3 g(blocks_for(n, S::EPB)), b(S::BLOCK);
    if (kind == av_kind_adv) hipLaunchKernelGGL((k_av_write_face<ND, RS, av_kind_adv>), g, b, 0, l.stream, l.m, o, cls);
    else if (kind == av_kind_sav) hipLaunchKernelGGL((k_av_write_face<ND, RS, av_kind_sav>), g, b, 0, l.stream, l.m, o, cls);
    else hipLaunchKernelGGL((k_av_write_face<ND, RS, av_kind_fta>), g, b, 0, l.stream, l.m, o, cls);
    return hipGetLastError();
  }
  static hipError_t av_neighbor(const Launch& l, int cls, int kind)
  {
    const int n = l.m.n_con[cls];
    if (n == 0) return hipSuccess;
    dim3 g(blocks_for(size_t(n) * S::NFQ, 256)), b(256);
    const Av_consts<RS> c = av_consts(l, RS);
    if (cls == HX_DEFORMED) {
      if (kind == av_kind_adv) hipLaunchKernelGGL((k_av_neighbor<ND, RS, true, av_kind_adv>), g, b, 0, l.stream, l.m, c);
      else if (kind == av_kind_sav) hipLaunchKernelGGL((k_av_neighbor<ND, RS, true, av_kind_sav>), g, b, 0, l.stream, l.m, c);
      else hipLaunchKernelGGL((k_av_neighbor<ND, RS, true, av_kind_fta>), g, b, 0, l.stream, l.m, c);
    } else {
      if (kind == av_kind_adv) hipLaunchKernelGGL((k_av_neighbor<ND, RS, false, av_kind_adv>), g, b, 0, l.stream, l.m, c);
      else if (kind == av_kind_sav) hipLaunchKernelGGL((k_av_neighbor<ND, RS, false, av_kind_sav>), g, b, 0, l.stream, l.m, c);
      else hipLaunchKernelGGL((k_av_neighbor<ND, RS, false, av_kind_fta>), g, b, 0, l.stream, l.m, c);
    }
    return hipGetLastError();
  }
  // project_av_rs > 0: the Legendre projection of degree project_av_rs - 1 into forcing[0] after the update
  static hipError_t adv_local(const Launch& l, int cls, double coef, int stage, int cr, double advect_length,
                              int project_av_rs)
  {
    const int n = l.m.n_elem[cls];
    if (n == 0) return hipSuccess;
    const Eops<RS, double> o = make_eops<RS, double>(hops(l));
    const Av_consts<RS> c = av_consts(l, project_av_rs > 0 ? project_av_rs : RS);
    const int project = project_av_rs > 0;
    dim3 g(blocks_for(n, S::EPB)), b(S::BLOCK);
    if (cls == HX_DEFORMED)
      hipLaunchKernelGGL((k_adv_local<ND, RS, true>), g, b, 0, l.stream, l.m, o, c, coef, stage, cr, advect_length, project);
    else
      hipLaunchKernelGGL((k_adv_local<ND, RS, false>), g, b, 0, l.stream, l.m, o, c, coef, stage, cr, advect_length, project);
    return hipGetLastError();
  }
  template <int KIND>
  static void diff_local_as(const Launch& l, int cls, double dt, double diff_time)
  {
    const Eops<RS, double> o = make_eops<RS, double>(hops(l));
    dim3 g(blocks_for(l.m.n_elem[cls], S::EPB)), b(S::BLOCK);
    if (cls == HX_DEFORMED) hipLaunchKernelGGL((k_sav_local<ND, RS, true, KIND>), g, b, 0, l.stream, l.m, o, dt, diff_time);
    else hipLaunchKernelGGL((k_sav_local<ND, RS, false, KIND>), g, b, 0, l.stream, l.m, o, dt, diff_time);
  }
  // Local of a diffusion-only PDE: Smooth_art_visc or Fix_therm_admis (kind)
  static hipError_t diff_local(const Launch& l, int cls, int kind, double dt, double diff_time)
  {
    if (l.m.n_elem[cls] == 0) return hipSuccess;
    if (kind == av_kind_sav) diff_local_as<av_kind_sav>(l, cls, dt, diff_time);
    else diff_local_as<av_kind_fta>(l, cls, dt, diff_time);
    return hipGetLastError();
  }
  static hipError_t diff_neighbor_reconcile(const Launch& l, int cls, int kind)
  {
    const int n = l.m.n_con[cls];
    if (n == 0) return hipSuccess;
    dim3 g(blocks_for(size_t(n) * S::NFQ, 256)), b(256);
    if (kind == av_kind_sav) {
      if (cls == HX_DEFORMED) hipLaunchKernelGGL((k_neighbor_reconcile<ND, RS, true, 3>), g, b, 0, l.stream, l.m);
      else hipLaunchKernelGGL((k_neighbor_reconcile<ND, RS, false, 3>), g, b, 0, l.stream, l.m);
    } else {
      if (cls == HX_DEFORMED) hipLaunchKernelGGL((k_neighbor_reconcile<ND, RS, true>), g, b, 0, l.stream, l.m);
      else hipLaunchKernelGGL((k_neighbor_reconcile<ND, RS, false>), g, b, 0, l.stream, l.m);
    }
    return hipGetLastError();
  }
  template <int KIND>
  static void diff_reconcile_as(const Launch& l, int cls, double dt, double diff_time, double cheby)
  {
    const Eops<RS, double> o = make_eops<RS, double>(hops(l));
    dim3 g(blocks_for(l.m.n_elem[cls], S::EPB)), b(S::BLOCK);
    if (cls == HX_DEFORMED)
      hipLaunchKernelGGL((k_sav_reconcile<ND, RS, true, KIND>), g, b, 0, l.stream, l.m, o, dt, diff_time, cheby);
    else hipLaunchKernelGGL((k_sav_reconcile<ND, RS, false, KIND>), g, b, 0, l.stream, l.m, o, dt, diff_time, cheby);
  }
  static hipError_t diff_reconcile(const Launch& l, int cls, int kind, double dt, double diff_time, double cheby)
  {
    if (l.m.n_elem[cls] == 0) return hipSuccess;
    if (kind == av_kind_sav) diff_reconcile_as<av_kind_sav>(l, cls, dt, diff_time, cheby);
    else diff_reconcile_as<av_kind_fta>(l, cls, dt, diff_time, cheby);
    return hipGetLastError();
  }
  // Solver::is_admissible over both classes and the hanging faces; flags[2] must be zero beforehand
  static hipError_t admissible(const Launch& l, int* flags)
  {
    for (int cls = 0; cls < 2; ++cls)
      if (l.m.n_elem[cls])
        hipLaunchKernelGGL((k_admissible<ND, RS>), dim3(blocks_for(l.m.n_elem[cls], S::EPB)), dim3(S::BLOCK), 0,
                           l.stream, l.m, cls, flags);
    if (l.m.n_ref)
      hipLaunchKernelGGL((k_admissible_ref<ND, RS>), dim3(blocks_for(size_t(l.m.n_ref) * (1 << (ND - 1)) * S::NFQ, 256)),
                         dim3(256), 0, l.stream, l.m, flags);
    return hipGetLastError();
  }
  // Solver::calc_jacobian pieces (hx_metrics_kernels.hpp)
  static hipError_t set_jacobian(const Launch& l, const double* vert, const double* adj, double* vtss,
                                 const double* node, const double* diff, const double* bound)
  {
    Metric_consts<RS> mc;
    for (int i = 0; i < RS; ++i) {
      mc.node[i] = node[i];
      for (int j = 0; j < RS; ++j) mc.diff[i][j] = diff[i * RS + j];
      mc.bound[0][i] = bound[i];
      mc.bound[1][i] = bound[RS + i];
    }
    if (l.m.n_elem[HX_DEFORMED])
      hipLaunchKernelGGL((k_set_jacobian<ND, RS>), dim3(blocks_for(l.m.n_elem[HX_DEFORMED], S::EPB)), dim3(S::BLOCK), 0,
                         l.stream, l.m, mc, vert, adj, vtss);
    if (l.m.n_elem[HX_CARTESIAN])
      hipLaunchKernelGGL((k_cartesian_jacobian<ND, RS>), dim3(blocks_for(size_t(l.m.n_elem[HX_CARTESIAN]) * 2 * ND * S::NFQ, 256)),
                         dim3(256), 0, l.stream, l.m, vtss);
    return hipGetLastError();
  }
  static hipError_t connection_normals(const Launch& l)
  {
    if (l.m.n_bc_con)
      hipLaunchKernelGGL((k_bc_copy_normal<ND, RS>), dim3(blocks_for(size_t(l.m.n_bc_con) * S::NFQ, 256)), dim3(256), 0,
                         l.stream, l.m);
    if (l.m.n_con[HX_DEFORMED])
      hipLaunchKernelGGL((k_connection_normals<ND, RS>), dim3(blocks_for(size_t(l.m.n_con[HX_DEFORMED]) * S::NFQ, 256)),
                         dim3(256), 0, l.stream, l.m);
    return hipGetLastError();
  }
  static hipError_t hanging_match(const Launch& l, const int* hanging, int n_hanging, double* vtss, double* vals, int mode)
  {
    if (n_hanging == 0) return hipSuccess;
    hipLaunchKernelGGL((k_hanging_match<ND>), dim3(blocks_for(n_hanging, 256)), dim3(256), 0, l.stream, hanging,
                       n_hanging, l.m.n_elem[HX_CARTESIAN], vtss, vals, mode);
    return hipGetLastError();
  }
  static hipError_t fix_av(const Launch& l, int mode)
  {
    for (int cls = 0; cls < 2; ++cls)
      if (l.m.n_elem[cls])
        hipLaunchKernelGGL((k_fix_av<ND, RS>), dim3(blocks_for(size_t(l.m.n_elem[cls]) * S::NQ, 256)), dim3(256), 0,
                           l.stream, l.m, cls, mode);
    return hipGetLastError();
  }
  static hipError_t av_max_dt(const Launch& l, int cls, int smooth, int is_local, double cc, double cd)
  {
    const int n = l.m.n_elem[cls];
    if (n == 0) return hipSuccess;
    const Av_consts<RS> c = av_consts(l, RS);
    dim3 g(std::min(blocks_for(size_t(n) * S::NQ, 256), unsigned(max_dt_blocks))), b(256);
    if (smooth) hipLaunchKernelGGL((k_av_max_dt<ND, RS, true>), g, b, 0, l.stream, l.m, c, cls, is_local, cc, cd, l.scalar);
    else hipLaunchKernelGGL((k_av_max_dt<ND, RS, false>), g, b, 0, l.stream, l.m, c, cls, is_local, cc, cd, l.scalar);
    return hipGetLastError();
  }
  static hipError_t av_bcs(const Launch& l, int mode)
  {
    if (l.m.n_bc_con == 0) return hipSuccess;
    hipLaunchKernelGGL((k_av_bcs<ND, RS>), dim3(blocks_for(size_t(l.m.n_bc_con) * S::NFQ, 256)), dim3(256), 0, l.stream,
                       l.m, mode);
    return hipGetLastError();
  }
  static hipError_t av_points(const Launch& l, int cls, int what, int av_rs, double mult, double us_max, double* resid)
  {
    const int n = l.m.n_elem[cls];
    if (n == 0) return hipSuccess;
    const Av_consts<RS> c = av_consts(l, av_rs);
    constexpr unsigned av_points_blocks = 2048;
    hipLaunchKernelGGL((k_av_points<ND, RS>), dim3(std::min(blocks_for(size_t(n) * S::NQ, 256), av_points_blocks)),
                       dim3(256), 0, l.stream, l.m, c, cls, what, mult, us_max, resid);
    return hipGetLastError();
  }
  // Stab_art_visc into per-element vertex slots (both classes: the deformed class after the Cartesian one), and the
  // vertex values interpolated into the Laplacian artificial viscosity
  static hipError_t stab_indicator(const Launch& l, int cls, double char_speed, double* vert)
  {
    const int n = l.m.n_elem[cls];
    if (n == 0) return hipSuccess;
    const Av_consts<RS> c = av_consts(l, RS);
    double* out = vert + (cls == HX_DEFORMED ? size_t(l.m.n_elem[HX_CARTESIAN]) << ND : 0);
    hipLaunchKernelGGL((k_stab_indicator<ND, RS>), dim3(blocks_for(n, S::EPB)), dim3(S::BLOCK), 0, l.stream, l.m, c, cls,
                       char_speed, out);
    return hipGetLastError();
  }
  static hipError_t vertex_interp_lap(const Launch& l, int cls, const double* vert)
  {
    const int n = l.m.n_elem[cls];
    if (n == 0) return hipSuccess;
    const Av_consts<RS> c = av_consts(l, RS);
    const double* in = vert + (cls == HX_DEFORMED ? size_t(l.m.n_elem[HX_CARTESIAN]) << ND : 0);
    hipLaunchKernelGGL((k_vertex_interp_lap<ND, RS>), dim3(blocks_for(size_t(n) * S::NQ, 256)), dim3(256), 0, l.stream,
                       l.m, c, cls, in);
    return hipGetLastError();
  }
  static const Kernel_set* set()
  {
    static const Kernel_set s{&write_face,  &neighbor,      &local,     &neighbor_reconcile, &reconcile_ldg_flux,
                              &max_dt,      &prolong,       &restrict_, &state_bcs,          &flux_bcs,
                              &av_write_face, &av_neighbor, &adv_local, &diff_local,         &diff_neighbor_reconcile,
                              &diff_reconcile, &av_max_dt,  &av_bcs,    &av_points,          &max_dt_cases,
                              &admissible,    &fix_av,      &set_jacobian, &connection_normals, &hanging_match,
                              &stab_indicator, &vertex_interp_lap, &mortar_restrict};
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
