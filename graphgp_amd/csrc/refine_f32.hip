// refine_f32.hip -- FP32 configurations of the fused refinement kernel.
//   k+1 <= 12 : one lane per point;  <= 20 : two lanes;  <= 32 : four lanes;  <= 40 : eight lanes.
#include "refine_launch.hpp"

namespace ggp {

extern template int launch_config<float, 24, 4, 3>(RefineArgs<float>, hipStream_t, size_t, int);
extern template int launch_config<float, 28, 4, 3>(RefineArgs<float>, hipStream_t, size_t, int);
extern template int launch_config<float, 32, 4, 3>(RefineArgs<float>, hipStream_t, size_t, int);
extern template int launch_config<float, 40, 8, 3>(RefineArgs<float>, hipStream_t, size_t, int);

template <>
int launch_refine<float>(const RefineArgs<float>& a, int kk, hipStream_t s, size_t lds, int grid) {
  if (a.d <= 3) {
    if (kk <= 4) return launch_config<float, 4, 1, 3>(a, s, lds, grid);
    if (kk <= 8) return launch_config<float, 8, 1, 3>(a, s, lds, grid);
    if (kk <= 12) return launch_config<float, 12, 1, 3>(a, s, lds, grid);
    if (kk <= 16) return launch_config<float, 16, 2, 3>(a, s, lds, grid);
    if (kk <= 20) return launch_config<float, 20, 2, 3>(a, s, lds, grid);
    if (kk <= 24) return launch_config<float, 24, 4, 3>(a, s, lds, grid);
    if (kk <= 28) return launch_config<float, 28, 4, 3>(a, s, lds, grid);
    if (kk <= 32) return launch_config<float, 32, 4, 3>(a, s, lds, grid);
    if (kk <= 40) return launch_config<float, 40, 8, 3>(a, s, lds, grid);
  } else if (a.d <= 8) {
    if (kk <= 8) return launch_config<float, 8, 1, 8>(a, s, lds, grid);
    if (kk <= 16) return launch_config<float, 16, 2, 8>(a, s, lds, grid);
  }
  return kUnsupported;
}

}  // namespace ggp
