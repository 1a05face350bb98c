// Device-resident bloom filters of predicate transfer:
//   ArrowBloomFilter  blocked filter (fpdb-executor/src/physical/bloomfilter/ArrowBloomFilter.cpp)
//   BloomFilter       legacy universal-hash filter (.../bloomfilter/BloomFilter.cpp)
// both behind BloomFilterBase (.../bloomfilter/BloomFilterBase.h:14-40).
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "Expected.h"
#include "TupleSet.h"

namespace fpdb::amd {

enum BloomFilterType { BLOOM_FILTER, ARROW_BLOOM_FILTER };

class BloomFilterBase {
 public:
  BloomFilterBase(BloomFilterType type, int64_t capacity, bool valid)
      : type_(type), capacity_(capacity), valid_(valid) {}
  virtual ~BloomFilterBase() = default;
  BloomFilterType getType() const { return type_; }
  int64_t capacity() const { return capacity_; }
  bool valid() const { return valid_; }
  virtual std::string toJson() const = 0;

 protected:
  BloomFilterType type_;
  int64_t capacity_;
  bool valid_;
};

class ArrowBloomFilter : public BloomFilterBase {
 public:
  ArrowBloomFilter(int64_t capacity, std::vector<std::string> columnNames);
  static std::shared_ptr<ArrowBloomFilter> make(int64_t capacity, const std::vector<std::string>& columnNames);

  // Hashes the key columns of every row and inserts them
  // (ArrowBloomFilter::build, ArrowBloomFilter.cpp:29-75).
  Expected<void> build(const std::shared_ptr<TupleSet>& tupleSet);
  // ORs the rows of another batch into the filter.
  Expected<void> insert(const std::shared_ptr<TupleSet>& tupleSet);

  int logNumBlocks() const { return logNumBlocks_; }
  const uint64_t* deviceBlocks() const { return static_cast<const uint64_t*>(blocks_->data()); }
  const std::vector<std::string>& columnNames() const { return columnNames_; }
  std::vector<uint64_t> blocksToHost() const;
  // Masks (136 bytes) + bit array, the two record batches of
  // makeBitmapRecordBatches (ArrowBloomFilter.cpp:77-116), as raw bytes.
  std::vector<uint8_t> masksBytes() const;
  std::string toJson() const override;

 private:
  std::vector<std::string> columnNames_;
  int logNumBlocks_;
  std::shared_ptr<DeviceBuffer> blocks_;
};

class BloomFilter : public BloomFilterBase {
 public:
  BloomFilter(int64_t capacity, double falsePositiveRate);
  static std::shared_ptr<BloomFilter> make(int64_t capacity, double falsePositiveRate);

  // k, m, hash functions and a zeroed bit array (BloomFilter::init, BloomFilter.cpp:55-60).
  void init(uint32_t seed);
  Expected<void> build(const std::shared_ptr<TupleSet>& tupleSet, const std::vector<std::string>& columnNames);
  // OR of two filters with the same capacity and rate (BloomFilter::merge, BloomFilter.cpp:101-121).
  Expected<void> merge(const std::shared_ptr<BloomFilter>& other);

  double falsePositiveRate() const { return fpr_; }
  int64_t numHashFunctions() const { return (int64_t)hashFunctions_.size(); }
  int64_t numBits() const { return numBits_; }
  const std::vector<fpdb_universal_hash>& hashFunctions() const { return hashFunctions_; }
  const int64_t* deviceWords() const { return static_cast<const int64_t*>(words_->data()); }
  std::vector<int64_t> getBitArray() const;
  std::string toJson() const override;

 private:
  double fpr_;
  int64_t numBits_ = 0;
  std::vector<fpdb_universal_hash> hashFunctions_;
  std::shared_ptr<DeviceBuffer> words_;
};

// Resolves key column names to columns of a TupleSet.
Expected<std::vector<fpdb_column>> keyColumns(const std::shared_ptr<TupleSet>& ts,
                                              const std::vector<std::string>& names);

}  // namespace fpdb::amd
