// adjoint_k_f32_inv.hip -- the register adjoint kernels for float, mode kInverse, in a file of their own
// so that the instantiations compile in parallel.
#include "adjoint_impl.hpp"

namespace ggp {
GGP_ADJOINT_INSTANTIATE(float, kInverse)
}  // namespace ggp
