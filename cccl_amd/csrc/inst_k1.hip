// Kernel instantiations for 1-byte keys.
#include "inst_impl.hpp"

namespace cccl_amd {

const KeyWidthTable& kernels_k1()
{
  static const KeyWidthTable table = inst::make_table<uint8_t>();
  return table;
}

} // namespace cccl_amd
