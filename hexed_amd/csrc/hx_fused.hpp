// Fused launches of an Euler stage over a group of meshes (hx_group): one launch of the state boundary conditions, one
// of Neighbor and one of Local cover every mesh of the group, both element classes and every row size of the set RSS,
// where the per-mesh path launches each kernel once per mesh and class. The stages fold the face fluxes into the Local
// (face_flux), as the per-mesh batch stage does, so Neighbor runs only on connections touching a hanging-face mortar.
//
// Work is cut into tiles -- one workgroup's worth of the per-class kernel (256 face items, or the EPB elements of an
// Euler Local workgroup) -- listed per (mesh, class) segment in a plan. A workgroup finds its tile's segment, and runs
// the same tile function the per-class kernel runs (local_euler_tile, neighbor_at, state_bcs_at) with the row size as a
// template argument, so every result is bit-identical to the per-mesh path. With a persistent grid each workgroup walks
// several tiles, strided over the grid or as one contiguous chunk (the launch shape the autotuner picks).
#pragma once

#include <type_traits>

#include "hx_kernels.hpp"

namespace hx {

// calls f(std::integral_constant<int, R>) for the R of RS... equal to rs
template <int R0, int... REST, class F>
__device__ __forceinline__ void dispatch_rs(int rs, F&& f)
{
  if (rs == R0) {
    f(std::integral_constant<int, R0>{});
  } else {
    if constexpr (sizeof...(REST) > 0) dispatch_rs<REST...>(rs, f);
  }
}

// the segment holding `tile` (plans have few segments; the scan is uniform across the workgroup)
__device__ __forceinline__ int fused_segment(const Fused_plan* __restrict__ p, int tile)
{
  int s = 0;
  while (s + 1 < p->n_seg && tile >= p->first[s + 1]) ++s;
  return s;
}

// the tiles of this workgroup: strided over the grid (order 0) or one contiguous chunk (order 1)
// PERSIST false: one tile per workgroup (the grid is the tile count), without a loop. A separate instantiation, because
// the compiler hoists the tiles' loop-invariant constants (operator tiles of every row size) out of a loop and holds
// them all in registers.
template <bool PERSIST, class F>
__device__ __forceinline__ void for_tiles(int n_tiles, int order, int chunk, F&& f)
{
  if constexpr (!PERSIST) {
    f(int(blockIdx.x));
  } else if (order == 0) {
    for (int t = blockIdx.x; t < n_tiles; t += gridDim.x) f(t);
  } else {
    const int a = blockIdx.x * chunk;
    const int b = a + chunk < n_tiles ? a + chunk : n_tiles;
    for (int t = a; t < b; ++t) f(t);
  }
}

// The operator constants of every row size of the set, a by-value kernel argument as in the per-class kernels (the
// compiler reloads them from the argument segment where they are used instead of holding them in registers)
template <int... RSS>
struct Eops_set;
template <>
struct Eops_set<>
{
};
template <int R0, int... REST>
struct Eops_set<R0, REST...>
{
  Eops<R0, double> head;
  Eops_set<REST...> tail;
};
template <int RS, int R0, int... REST>
__device__ __forceinline__ const Eops<RS, double>& eops_of(const Eops_set<R0, REST...>& s)
{
  if constexpr (RS == R0) return s.head;
  else return eops_of<RS>(s.tail);
}
template <int R0, int... REST>
Eops_set<R0, REST...> make_eops_set(const void* const* host_ops)
{
  Eops_set<R0, REST...> s;
  s.head = make_eops<R0, double>(*static_cast<const Ops<R0>*>(host_ops[R0]));
  if constexpr (sizeof...(REST) > 0) s.tail = make_eops_set<REST...>(host_ops);
  return s;
}

template <int ND, int... RSS>
__global__ __launch_bounds__(256) void k_fused_state_bcs(const Fused_member* __restrict__ mem,
                                                        const Fused_plan* __restrict__ plan, int order, int chunk)
{
  for_tiles<false>(plan->n_tiles, order, chunk, [&](int tile) {
    const int s = fused_segment(plan, tile);
    const Dev_mesh& m = mem[plan->member[s]].m;
    const int gid = (tile - plan->first[s]) * 256 + int(threadIdx.x);
    dispatch_rs<RSS...>(plan->rs[s], [&](auto rs_c) { state_bcs_at<ND, decltype(rs_c)::value>(m, gid); });
  });
}

// with face-flux folding only the connections touching a hanging-face mortar (Dev_mesh::mortar_con) run Neighbor
template <int ND, int... RSS>
__global__ __launch_bounds__(256) void k_fused_neighbor(const Fused_member* __restrict__ mem,
                                                       const Fused_plan* __restrict__ plan, int order, int chunk)
{
  for_tiles<false>(plan->n_tiles, order, chunk, [&](int tile) {
    const int s = fused_segment(plan, tile);
    const Dev_mesh& m = mem[plan->member[s]].m;
    const int gid = (tile - plan->first[s]) * 256 + int(threadIdx.x);
    const int cls = plan->cls[s];
    dispatch_rs<RSS...>(plan->rs[s], [&](auto rs_c) {
      constexpr int RS = decltype(rs_c)::value;
      constexpr int NFQ = Shape<ND, RS>::NFQ;
      const int i = gid / NFQ, f = gid % NFQ;
      if (i >= m.n_mortar_con[cls]) return;
      if (cls == HX_DEFORMED) neighbor_con<ND, RS, true, false>(m, m.mortar_con[cls][i], f);
      else neighbor_con<ND, RS, false, false>(m, m.mortar_con[cls][i], f);
    });
  });
}

// the FP32-MFMA form's operand table at the start of the fused Local's LDS (the largest of the set's row sizes),
// in doubles so that the tile arrays after it stay 8-byte aligned
template <class P, int... RSS>
constexpr int fused_table_doubles()
{
  int n = 0;
  if constexpr (P::mfma)
    for (int x : {4 * Eops<RSS, double>::MF_STRIDE...}) n = x > n ? x : n;
  return (n + 1) / 2;
}

// largest LDS footprint and smallest register target of the Euler Local tiles the fused kernel may run (the operand
// table included)
template <int ND, class P, bool WITH_DEF, int... RSS>
constexpr int fused_local_lds()
{
  int b = 0;
  for (int x : {euler_lds_bytes<ND, RSS, false, P>()...}) b = x > b ? x : b;
  if constexpr (WITH_DEF)
    for (int x : {euler_lds_bytes<ND, RSS, true, euler_form<P, true>>()...}) b = x > b ? x : b;
  return b + 8 * fused_table_doubles<P, RSS...>();
}
// The FP32-MFMA form with one tile per workgroup takes 5 waves per SIMD where the LDS allows them, although the row
// sizes whose tiles give each thread two face points (4 and 5 in 3D) then spill a few registers: the larger row sizes
// carry most of the work (measured on a batch of row sizes 4 to 6: 2.71 -> 2.59 ms per step). Otherwise the target is
// the smallest of the per-class targets (the FP64 form, which the LDS holds to 4 workgroups per compute unit, only
// spills at 5; a persistent grid, whose tile loop holds more registers, spills far more: 1.9 -> 3.0 ms per launch).
template <int ND, class P, bool WITH_DEF, bool PERSIST, int... RSS>
constexpr int fused_local_waves()
{
  int w = 8;
  for (int x : {euler_waves<ND, RSS, false, P, true>()...}) w = x < w ? x : w;
  if constexpr (WITH_DEF)
    for (int x : {euler_waves<ND, RSS, true, euler_form<P, true>, true>()...}) w = x < w ? x : w;
  // workgroups of 4 waves (one per SIMD) that the LDS holds per compute unit
  if constexpr (P::mfma && !PERSIST)
    if (163840 / fused_local_lds<ND, P, WITH_DEF, RSS...>() >= 5 && w < 5) w = 5;
  return w;
}

template <int ND, class P, bool WITH_DEF, bool PERSIST, int... RSS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(fused_local_waves<ND, P, WITH_DEF, PERSIST, RSS...>())))
void k_fused_local_euler(const Fused_member* __restrict__ mem, const Fused_plan* __restrict__ plan, int order, int chunk,
                         Eops_set<RSS...> ops, double coef, const double* __restrict__ dtp, int stage,
                         int compute_residual, int use_filter)
{
  extern __shared__ double fused_lds_all[];
  float* mtab = reinterpret_cast<float*>(fused_lds_all);
  double* fused_lds = fused_lds_all + fused_table_doubles<P, RSS...>();
  for_tiles<PERSIST>(plan->n_tiles, order, chunk, [&](int tile) {
    const int s = fused_segment(plan, tile);
    const Fused_member& mb = mem[plan->member[s]];
    const int blk = tile - plan->first[s];
    const bool def = plan->cls[s] == HX_DEFORMED;
    dispatch_rs<RSS...>(plan->rs[s], [&](auto rs_c) {
      constexpr int RS = decltype(rs_c)::value;
      using T = typename P::T;
      const Eops<RS, double>& o = eops_of<RS>(ops);
      const Eops<RS, double>* od = static_cast<const Eops<RS, double>*>(mb.eops);
      const Ops<RS>* gops = static_cast<const Ops<RS>*>(mb.ops);
      if (def) {
        if constexpr (WITH_DEF) {
          using PD = euler_form<P, true>;
          using TD = typename PD::T;
          constexpr int tb = Shape<ND, RS>::EPB * int(sizeof(TD)) * Euler_lds<ND, RS, true, PD>::PER_T;
          TD* lt = reinterpret_cast<TD*>(fused_lds);
          double* ld = fused_lds + (tb + 7) / 8;
          local_euler_tile<ND, RS, true, PD, true>(mb.m, blk, lt, ld, mtab, o, od, coef, dtp, stage,
                                                   compute_residual, use_filter, mb.fuse, gops);
        }
      } else {
        constexpr int tb = Shape<ND, RS>::EPB * int(sizeof(T)) * Euler_lds<ND, RS, false, P>::PER_T;
        T* lt = reinterpret_cast<T*>(fused_lds);
        double* ld = fused_lds + (tb + 7) / 8;
        local_euler_tile<ND, RS, false, P, true>(mb.m, blk, lt, ld, mtab, o, od, coef, dtp, stage, compute_residual,
                                                 use_filter, mb.fuse, gops);
      }
    });
    // the next tile reuses the LDS
    if constexpr (PERSIST) __syncthreads();
  });
}

// tiles of one (mesh, class) segment: 256 boundary or connection face items, or one Euler Local workgroup of elements
template <int ND, int RS>
inline int fused_tiles(int kind, const Dev_mesh& m, int cls)
{
  using S = Shape<ND, RS>;
  if (kind == fused_state_bcs) return int((size_t(m.n_bc_con) * S::NFQ + 255) / 256);
  if (kind == fused_neighbor) return int((size_t(m.n_mortar_con[cls]) * S::NFQ + 255) / 256);
  return (m.n_elem[cls] + S::EPB - 1) / S::EPB;
}

template <int ND, int... RSS>
struct Fused_launchers
{
  static bool supports(int rs) { return ((rs == RSS) || ...); }
  static hipError_t state_bcs(const Fused_launch& l)
  {
    if (!l.n_tiles) return hipSuccess;
    hipLaunchKernelGGL((k_fused_state_bcs<ND, RSS...>), dim3(l.grid), dim3(256), 0, l.stream, l.mem, l.plan, l.order,
                       l.chunk);
    return hipGetLastError();
  }
  static hipError_t neighbor(const Fused_launch& l)
  {
    if (!l.n_tiles) return hipSuccess;
    hipLaunchKernelGGL((k_fused_neighbor<ND, RSS...>), dim3(l.grid), dim3(256), 0, l.stream, l.mem, l.plan, l.order,
                       l.chunk);
    return hipGetLastError();
  }
  template <class P, bool WITH_DEF>
  static hipError_t local_as(const Fused_launch& l, double coef, const double* dtp, int stage, int cr, int uf)
  {
    constexpr int lds = fused_local_lds<ND, P, WITH_DEF, RSS...>();
    const Eops_set<RSS...> ops = make_eops_set<RSS...>(l.host_ops);
    if (l.order == 2)
      hipLaunchKernelGGL((k_fused_local_euler<ND, P, WITH_DEF, false, RSS...>), dim3(l.grid), dim3(256), lds, l.stream,
                         l.mem, l.plan, l.order, l.chunk, ops, coef, dtp, stage, cr, uf);
    else
      hipLaunchKernelGGL((k_fused_local_euler<ND, P, WITH_DEF, true, RSS...>), dim3(l.grid), dim3(256), lds, l.stream,
                         l.mem, l.plan, l.order, l.chunk, ops, coef, dtp, stage, cr, uf);
    return hipGetLastError();
  }
  static hipError_t local_euler(const Fused_launch& l, int precision, bool with_def, double coef, const double* dtp,
                                int stage, int cr, int uf)
  {
    if (!l.n_tiles) return hipSuccess;
    if (precision == HX_FP32_MFMA)
      return with_def ? local_as<Fp32_mfma_form, true>(l, coef, dtp, stage, cr, uf)
                      : local_as<Fp32_mfma_form, false>(l, coef, dtp, stage, cr, uf);
    return with_def ? local_as<Fp64_form, true>(l, coef, dtp, stage, cr, uf)
                    : local_as<Fp64_form, false>(l, coef, dtp, stage, cr, uf);
  }
  // tiles of one segment of each fused kernel
  static int tiles(int kind, const Dev_mesh& m, int cls)
  {
    int out = 0;
    ((m.rs == RSS ? (out = fused_tiles<ND, RSS>(kind, m, cls)) : 0), ...);
    return out;
  }
  static const Fused_set* set()
  {
    static const Fused_set s{&supports, &state_bcs, &neighbor, &local_euler, &tiles};
    return &s;
  }
};

}  // namespace hx
