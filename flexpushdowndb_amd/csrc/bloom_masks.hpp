// The blocked bloom filter's mask bit vector: 1024 overlapping 57-bit masks,
// each window holding 4 or 5 set bits.  These are the 136 bytes that
// BloomFilterMasks() (fpdb-tuple/src/arrow/exec/BloomFilter.cpp:16-79) produces
// from std::mt19937(seed_seq{0,0,0,0,0,0,0,0}) with libstdc++'s
// uniform_int_distribution<uint64_t>; the value is a fixed table of the data
// format, embedded so the product does not depend on the host C++ library.
#pragma once
#include <cstdint>

namespace fpdb {

constexpr int kMaskBitsPerMask = 57;
constexpr int kMaskLogNum = 10;
constexpr int kMaskNum = 1 << kMaskLogNum;
constexpr int kMaskTotalBytes = (kMaskNum + 64) / 8;   // 136

constexpr uint8_t kBloomMaskBytes[kMaskTotalBytes] = {
#include "bloom_masks.inc"
};

struct ExpandedMasks {
  uint64_t m[kMaskNum];
};

// mask(i): little-endian unaligned 64-bit load at byte i/8, >> (i%8), 57 bits
// (BloomFilter.h:52-61).
constexpr ExpandedMasks expand_masks() {
  ExpandedMasks t{};
  for (int i = 0; i < kMaskNum; ++i) {
    uint64_t w = 0;
    for (int b = 0; b < 8; ++b) {
      int idx = i / 8 + b;
      uint64_t byte = idx < kMaskTotalBytes ? kBloomMaskBytes[idx] : 0;
      w |= byte << (8 * b);
    }
    t.m[i] = (w >> (i % 8)) & ((1ULL << kMaskBitsPerMask) - 1);
  }
  return t;
}

}  // namespace fpdb
