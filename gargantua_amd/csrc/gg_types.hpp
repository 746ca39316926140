// gg_types.hpp -- plain value types shared by device code and host-compiled
// checks of the same sources.
#pragma once

#ifdef __HIPCC__
#define GG_HD __host__ __device__
#else
#define GG_HD
#endif

namespace gg {

struct d4 { double x, y, z, w; };
struct f4 { float x, y, z, w; };

}  // namespace gg
