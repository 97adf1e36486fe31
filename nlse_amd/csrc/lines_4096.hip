// Kernels of line length 4096.
#include "lines.hpp"

namespace nlse {
LineOps make_ops_4096() { return make_ops<4096>(); }
}  // namespace nlse
