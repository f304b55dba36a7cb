// Line-solve kernels instantiated for float (see lines_impl.h).
#include "lines_impl.h"

namespace x3d {
template int launch_line<float>(const LineCall<float> &, hipStream_t);
}  // namespace x3d
