// Kernels of line length 512.
#include "lines.hpp"

namespace nlse {
LineOps make_ops_512() { return make_ops<512>(); }
}  // namespace nlse
