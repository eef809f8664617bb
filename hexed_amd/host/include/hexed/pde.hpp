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

}  // namespace hexed::pde
#endif
