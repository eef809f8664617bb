// Transport coefficients (Transport_model.hpp:14-56 in the hexed tree): a constant or Sutherland's law.
#ifndef HEXED_AMD_TRANSPORT_MODEL_HPP_
#define HEXED_AMD_TRANSPORT_MODEL_HPP_

#include "../../../../include/hexed_amd.h"

namespace hexed
{

class Transport_model
{
  hx_transport t;
  Transport_model(double cv, double rv, double rt, double to, bool iv) : t{cv, rv, rt, to, int(iv)}, is_viscous{iv} {}

  public:
  const bool is_viscous;
  Transport_model(const Transport_model& o) : t{o.t}, is_viscous{o.is_viscous} {}
  static Transport_model inviscid() { return {0., 0., 1., 1., false}; }
  static Transport_model constant(double value) { return {value, 0., 1., 1., true}; }
  static Transport_model sutherland(double reference_value, double reference_temperature, double temperature_offset)
  {
    return {0., reference_value, reference_temperature, temperature_offset, true};
  }
  const hx_transport& abi() const { return t; }
};

inline const Transport_model inviscid = Transport_model::inviscid();

}  // namespace hexed
#endif
