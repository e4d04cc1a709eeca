// ibp_fetch_narrow.hip -- the fetch kernels for 2- and 1-byte elements (see ibp_fetch_launch.h).
#include "ibp_fetch_launch.h"

namespace ibp {

template int fetch_launch_width<uint16_t>(const FetchParams &, int, int, int, int, int, int, int, hipStream_t);
template int fetch_launch_width<uint8_t>(const FetchParams &, int, int, int, int, int, int, int, hipStream_t);

} // namespace ibp
