// refine_f32_big.hip -- the larger-k float configurations of the refinement kernel, compiled
// in their own unit so the build parallelises.
#include "refine_launch.hpp"

namespace ggp {

template int launch_config<float, 24, 4, 3>(RefineArgs<float>, hipStream_t, size_t, int);
template int launch_config<float, 28, 4, 3>(RefineArgs<float>, hipStream_t, size_t, int);
template int launch_config<float, 32, 4, 3>(RefineArgs<float>, hipStream_t, size_t, int);
template int launch_config<float, 40, 8, 3>(RefineArgs<float>, hipStream_t, size_t, int);

}  // namespace ggp
