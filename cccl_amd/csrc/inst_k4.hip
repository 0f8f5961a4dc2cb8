// Kernel instantiations for 4-byte keys.
#include "inst_impl.hpp"

namespace cccl_amd {

const KeyWidthTable& kernels_k4()
{
  static const KeyWidthTable table = inst::make_table<uint32_t>();
  return table;
}

} // namespace cccl_amd
