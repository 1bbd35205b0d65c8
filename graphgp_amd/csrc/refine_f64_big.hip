// refine_f64_big.hip -- the larger-k double configurations of the refinement kernel, compiled
// in their own unit so the build parallelises.
#include "refine_launch.hpp"

namespace ggp {

template int launch_config<double, 20, 4, 3>(RefineArgs<double>, hipStream_t, size_t, int);
template int launch_config<double, 24, 4, 3>(RefineArgs<double>, hipStream_t, size_t, int);
template int launch_config<double, 32, 8, 3>(RefineArgs<double>, hipStream_t, size_t, int);
template int launch_config<double, 40, 8, 3>(RefineArgs<double>, hipStream_t, size_t, int);

}  // namespace ggp
