// Kernels of line length 1024.
#include "lines.hpp"

namespace nlse {
LineOps make_ops_1024() { return make_ops<1024>(); }
}  // namespace nlse
