// ibp_fetch_u64.hip -- the fetch kernels for 8-byte elements (see ibp_fetch_launch.h).
#include "ibp_fetch_launch.h"

namespace ibp {

template int fetch_launch_width<uint64_t>(const FetchParams &, int, int, int, int, int, int, int, hipStream_t);

} // namespace ibp
