// Instantiates the DG kernels for n_dim = 1, row sizes 2..8.
#include "hx_kernels.hpp"
HX_DEFINE_KERNEL_SET(1)
