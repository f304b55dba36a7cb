// Line-solve kernels instantiated for double (see lines_impl.h).
#include "lines_impl.h"

namespace x3d {
template int launch_line<double>(const LineCall<double> &, hipStream_t);
}  // namespace x3d
