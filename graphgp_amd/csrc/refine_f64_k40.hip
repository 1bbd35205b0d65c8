// refine_f64_k40.hip -- the k+1 = 40 double configuration of the refinement kernel (eight
// lanes per point), compiled in a unit of its own so the build parallelises.
#include "refine_launch.hpp"

namespace ggp {

template int launch_config<double, 40, 8, 3>(RefineArgs<double>, hipStream_t, size_t, int);

}  // namespace ggp
