// adjoint_k_f32_ld.hip -- the register adjoint kernels for float, mode kLogdet, in a file of their own
// so that the instantiations compile in parallel.
#include "adjoint_impl.hpp"

namespace ggp {
GGP_ADJOINT_INSTANTIATE(float, kLogdet)
}  // namespace ggp
