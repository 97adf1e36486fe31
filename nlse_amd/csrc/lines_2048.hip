// Kernels of line length 2048.
#include "lines.hpp"

namespace nlse {
LineOps make_ops_2048() { return make_ops<2048>(); }
}  // namespace nlse
