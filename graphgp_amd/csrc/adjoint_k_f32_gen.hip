// adjoint_k_f32_gen.hip -- the register adjoint kernels for float, mode kGenerate, in a file of their own
// so that the instantiations compile in parallel.
#include "adjoint_impl.hpp"

namespace ggp {
GGP_ADJOINT_INSTANTIATE(float, kGenerate)
}  // namespace ggp
