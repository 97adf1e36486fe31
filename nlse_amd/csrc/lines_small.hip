// Kernels of line lengths 8, 16, 32, 64, 128, 256.
#include "lines.hpp"

namespace nlse {
LineOps make_ops_8() { return make_ops<8>(); }
LineOps make_ops_16() { return make_ops<16>(); }
LineOps make_ops_32() { return make_ops<32>(); }
LineOps make_ops_64() { return make_ops<64>(); }
LineOps make_ops_128() { return make_ops<128>(); }
LineOps make_ops_256() { return make_ops<256>(); }
}  // namespace nlse
