// refine_f32_k40.hip -- the k+1 = 40 float configuration of the refinement kernel (eight
// lanes per point), compiled in a unit of its own so the build parallelises.
#include "refine_launch.hpp"

namespace ggp {

template int launch_config<float, 40, 8, 3>(RefineArgs<float>, hipStream_t, size_t, int);

}  // namespace ggp
