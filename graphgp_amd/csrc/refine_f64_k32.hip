// refine_f64_k32.hip -- the k+1 = 32 double configuration of the refinement kernel (eight
// lanes per point), compiled in a unit of its own so the build parallelises.
#include "refine_launch.hpp"

namespace ggp {

template int launch_config<double, 32, 8, 3>(RefineArgs<double>, hipStream_t, size_t, int);

}  // namespace ggp
