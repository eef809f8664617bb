// PDE tags for Spatial<Pde_templ, is_deformed> (pde.hpp:27-259 in the hexed tree). They carry the constructor
// arguments of each PDE to the device kernels as an hx_pde; the flux formulas themselves live in the gfx950 kernels.
#ifndef HEXED_AMD_PDE_HPP_
#define HEXED_AMD_PDE_HPP_

#include "Transport_model.hpp"

namespace hexed::pde
{

template <bool visc = false>
class Navier_stokes
{
  public:
  Navier_stokes() = delete;
  template <int n_dim, int row_size>
  class Pde
  {
    public:
    static constexpr int kind = visc ? HX_PDE_NAVIER_STOKES : HX_PDE_EULER;
    static constexpr bool has_diffusion = visc;
    static constexpr bool has_convection = true;
    static constexpr int n_update = n_dim + 2;
    static constexpr int n_extrap = n_dim + 2;
    hx_pde abi;
    explicit Pde(Transport_model dynamic_visc = inviscid, Transport_model thermal_cond = inviscid)
    : abi{visc ? HX_PDE_NAVIER_STOKES : HX_PDE_EULER, dynamic_visc.abi(), thermal_cond.abi(), 0., 0., 0.}
    {}
  };
};

//! row_size scalars advected by the velocity held in the momentum slots (pde.hpp:265-349)
template <int n_dim, int row_size>
class Advection
{
  public:
  static constexpr int kind = HX_PDE_ADVECTION;
  static constexpr bool has_diffusion = false;
  static constexpr bool has_convection = true;
  static constexpr int n_update = row_size;
  static constexpr int n_extrap = n_dim + row_size;
  hx_pde abi;
  explicit Advection(double advect_length) : abi{HX_PDE_ADVECTION, inviscid.abi(), inviscid.abi(), advect_length, 0., 0.}
  {}
};

//! the diffusion chain that smooths the artificial-viscosity forcing (pde.hpp:355-431)
template <int n_dim, int row_size>
class Smooth_art_visc
{
  public:
  static constexpr int kind = HX_PDE_SMOOTH_AV;
  static constexpr bool has_diffusion = true;
  static constexpr bool has_convection = false;
  static constexpr int n_update = 3;
  static constexpr int n_extrap = 3;
  hx_pde abi;
  Smooth_art_visc(double diff_time, double cheby_step)
  : abi{HX_PDE_SMOOTH_AV, inviscid.abi(), inviscid.abi(), 0., diff_time, cheby_step}
  {}
};

//! unit diffusion of the conserved state that restores thermodynamic admissibility (pde.hpp:437-493)
template <int n_dim, int row_size>
class Fix_therm_admis
{
  public:
  static constexpr int kind = HX_PDE_FIX_ADMIS;
  static constexpr bool has_diffusion = true;
  static constexpr bool has_convection = false;
  static constexpr int n_update = n_dim + 2;
  static constexpr int n_extrap = n_dim + 2;
  hx_pde abi;
  Fix_therm_admis() : abi{HX_PDE_FIX_ADMIS, inviscid.abi(), inviscid.abi(), 0., 0., 0.} {}
};

//! the hx_pde of a PDE tag for the kernels that take no PDE arguments (Neighbor_reconcile, Prolong/Restrict_refined):
//! only the kind, which fixes the variables moved, matters
template <class P>
hx_pde kind_of()
{
  hx_pde p{};
  p.kind = P::kind;
  p.advect_length = p.diff_time = p.cheby_step = 1.;
  return p;
}

}  // namespace hexed::pde
#endif
