// ibp_fetch_u32.hip -- the fetch kernels for 4-byte elements (see ibp_fetch_launch.h).
#include "ibp_fetch_launch.h"

namespace ibp {

template int fetch_launch_width<uint32_t>(const FetchParams &, int, int, int, int, int, int, int, hipStream_t);

} // namespace ibp
