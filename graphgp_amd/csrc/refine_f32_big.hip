// refine_f32_big.hip -- the k+1 = 24 and 28 float configurations of the refinement kernel,
// compiled in a unit of its own so the build parallelises.
#include "refine_launch.hpp"

namespace ggp {

template int launch_config<float, 24, 4, 3>(RefineArgs<float>, hipStream_t, size_t, int);
template int launch_config<float, 28, 4, 3>(RefineArgs<float>, hipStream_t, size_t, int);

}  // namespace ggp
