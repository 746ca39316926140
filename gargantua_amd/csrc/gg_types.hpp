// gg_types.hpp -- plain value types shared by device code and host-compiled
// checks of the same sources.
#pragma once

#ifdef __HIPCC__
#define GG_HD __host__ __device__
#else
#define GG_HD
#endif

// Multiply-add contraction allowed in the enclosing block (the library is
// built with it off, which the forward integration needs to round like the
// reference); for the adjoint's own algebra, whose results are checked to a
// tolerance.  Other compilers of the host checks ignore it.
#ifdef __clang__
#define GG_CONTRACT_FAST _Pragma("clang fp contract(fast)")
#else
#define GG_CONTRACT_FAST
#endif

namespace gg {

struct d4 { double x, y, z, w; };
struct f4 { float x, y, z, w; };

}  // namespace gg
