// Operator-level constants and implementation switches, mirroring
// fpdb-executor/include/fpdb/executor/physical/Globals.h:19,47-55 and
// fpdb-tuple/include/fpdb/tuple/Globals.h:18.
#pragma once

#include <cstdint>

namespace fpdb::amd {

// rows an operator buffers before sending output (Globals.h:19)
inline int64_t DefaultBufferSize = 100000;
// bloom filters are not built over more input rows than this (Globals.h:55)
inline constexpr int64_t BLOOM_FILTER_MAX_INPUT_SIZE = 20000000;
// blocked (Arrow) bloom filter by default, legacy universal-hash filter otherwise (Globals.h:50)
inline bool USE_ARROW_BLOOM_FILTER_IMPL = true;
// seed for the legacy filter's universal-hash parameters (the reference draws
// them from std::random_device; here they are reproducible)
inline uint32_t LEGACY_BLOOM_FILTER_SEED = 0x5eed;

}  // namespace fpdb::amd
