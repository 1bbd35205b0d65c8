// adjoint_f64.hip -- FP64 instantiations of the refinement adjoint (own file so the
// FP32 and FP64 variants compile in parallel).
#include "adjoint_impl.hpp"

namespace ggp {

template <>
int launch_adjoint<double>(AdjointArgs<double> a, int kk, hipStream_t s, size_t lds, int grid) {
  if (a.d > 3) return kUnsupported;
  if (kk <= 8) return launch_adjoint_config<double, 8, 2, 3>(a, s, lds, grid);
  if (kk <= 12) return launch_adjoint_config<double, 12, 4, 3>(a, s, lds, grid);
  if (kk <= 16) return launch_adjoint_config<double, 16, 4, 3>(a, s, lds, grid);
  if (kk <= 24) return launch_adjoint_config<double, 24, 8, 3>(a, s, lds, grid);
  if (kk <= 32) return launch_adjoint_config<double, 32, 16, 3>(a, s, lds, grid);
  return kUnsupported;
}

}  // namespace ggp
