// Kernel instantiations for 8-byte keys.
#include "inst_impl.hpp"

namespace cccl_amd {

const KeyWidthTable& kernels_k8()
{
  static const KeyWidthTable table = inst::make_table<uint64_t>();
  return table;
}

} // namespace cccl_amd
