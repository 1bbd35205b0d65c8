// refine_f64_big.hip -- the k+1 = 20 and 24 double configurations of the refinement
// kernel, compiled in a unit of its own so the build parallelises.
#include "refine_launch.hpp"

namespace ggp {

template int launch_config<double, 20, 4, 3>(RefineArgs<double>, hipStream_t, size_t, int);
template int launch_config<double, 24, 4, 3>(RefineArgs<double>, hipStream_t, size_t, int);

}  // namespace ggp
