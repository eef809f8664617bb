// Instantiates the DG kernels for n_dim = 3, row sizes 2..8, and the fused launches for row sizes 4, 5 and 6.
#include "hx_kernels.hpp"
#include "hx_fused.hpp"
HX_DEFINE_KERNEL_SET(3)

namespace hx {
const Fused_set* fused_set_nd3() { return Fused_launchers<3, 4, 5, 6>::set(); }
}  // namespace hx
