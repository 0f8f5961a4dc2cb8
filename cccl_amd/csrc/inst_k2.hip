// Kernel instantiations for 2-byte keys.
#include "inst_impl.hpp"

namespace cccl_amd {

const KeyWidthTable& kernels_k2()
{
  static const KeyWidthTable table = inst::make_table<uint16_t>();
  return table;
}

} // namespace cccl_amd
