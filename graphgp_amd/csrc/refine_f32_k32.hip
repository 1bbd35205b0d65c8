// refine_f32_k32.hip -- the k+1 = 32 float configuration of the refinement kernel (four
// lanes per point), compiled in a unit of its own so the build parallelises.
#include "refine_launch.hpp"

namespace ggp {

template int launch_config<float, 32, 4, 3>(RefineArgs<float>, hipStream_t, size_t, int);

}  // namespace ggp
