// adjoint_k_f64_inv.hip -- the register adjoint kernels for double, mode kInverse, in a file of their own
// so that the instantiations compile in parallel.
#include "adjoint_impl.hpp"

namespace ggp {
GGP_ADJOINT_INSTANTIATE(double, kInverse)
}  // namespace ggp
