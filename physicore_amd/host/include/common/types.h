// Scalar types of the biofvm interface (reference: common/include/common/types.h:5-9).
#pragma once

#include <cstdint>

namespace physicore {

using real_t = double;
using index_t = std::uint64_t;
using sindex_t = std::int64_t;

} // namespace physicore
