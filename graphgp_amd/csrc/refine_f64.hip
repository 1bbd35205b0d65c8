// refine_f64.hip -- FP64 configurations of the fused refinement kernel.
//   k+1 <= 4 : one lane per point;  <= 12 : two lanes;  <= 24 : four lanes;  <= 40 : eight lanes.
#include "refine_launch.hpp"

namespace ggp {

extern template int launch_config<double, 20, 4, 3>(RefineArgs<double>, hipStream_t, size_t, int);
extern template int launch_config<double, 24, 4, 3>(RefineArgs<double>, hipStream_t, size_t, int);
extern template int launch_config<double, 32, 8, 3>(RefineArgs<double>, hipStream_t, size_t, int);
extern template int launch_config<double, 40, 8, 3>(RefineArgs<double>, hipStream_t, size_t, int);

template <>
int launch_refine<double>(const RefineArgs<double>& a, int kk, hipStream_t s, size_t lds, int grid) {
  if (a.d <= 3) {
    if (kk <= 4) return launch_config<double, 4, 1, 3>(a, s, lds, grid);
    if (kk <= 8) return launch_config<double, 8, 2, 3>(a, s, lds, grid);
    if (kk <= 12) return launch_config<double, 12, 2, 3>(a, s, lds, grid);
    if (kk <= 16) return launch_config<double, 16, 4, 3>(a, s, lds, grid);
    if (kk <= 20) return launch_config<double, 20, 4, 3>(a, s, lds, grid);
    if (kk <= 24) return launch_config<double, 24, 4, 3>(a, s, lds, grid);
    if (kk <= 32) return launch_config<double, 32, 8, 3>(a, s, lds, grid);
    if (kk <= 40) return launch_config<double, 40, 8, 3>(a, s, lds, grid);
  } else if (a.d <= 8) {
    if (kk <= 8) return launch_config<double, 8, 2, 8>(a, s, lds, grid);
    if (kk <= 16) return launch_config<double, 16, 4, 8>(a, s, lds, grid);
  }
  return kUnsupported;
}

}  // namespace ggp
