// adjoint_k_f64_gen.hip -- the register adjoint kernels for double, mode kGenerate, in a file of their own
// so that the instantiations compile in parallel.
#include "adjoint_impl.hpp"

namespace ggp {
GGP_ADJOINT_INSTANTIATE(double, kGenerate)
}  // namespace ggp
