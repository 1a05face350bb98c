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
#include "Json.h"
#include "TupleSet.h"

namespace fpdb::amd {

enum BloomFilterType { BLOOM_FILTER, ARROW_BLOOM_FILTER };

// One record batch of the bitmap wire form: a single column named "bitmap"
// (ArrowSerializer::BITMAP_FIELD_NAME, fpdb-tuple/include/fpdb/tuple/serialization/ArrowSerializer.h:68)
// of uint8 or int64 values, kept as its raw little-endian value bytes; the
// Arrow IPC framing around it is the transport's.
struct BitmapRecordBatch {
  std::string fieldName = "bitmap";
  fpdb_type type = FPDB_TYPE_UINT8;
  int64_t length = 0;                 // values
  std::vector<uint8_t> bytes;         // length * width bytes
};

// ArrowSerializer::bitmap_to_recordBatch / recordBatch_to_bitmap
// (fpdb-tuple/src/serialization/ArrowSerializer.cpp:330-365).
BitmapRecordBatch bitmapToRecordBatch(const std::vector<int64_t>& bitmap);
Expected<std::vector<int64_t>> recordBatchToBitmap(const BitmapRecordBatch& batch);

class BloomFilterBase {
 public:
  BloomFilterBase(BloomFilterType type, int64_t capacity, bool valid)
      : type_(type), capacity_(capacity), valid_(valid) {}
  virtual ~BloomFilterBase() = default;
  BloomFilterType getType() const { return type_; }
  int64_t capacity() const { return capacity_; }
  bool valid() const { return valid_; }
  virtual std::string toJson() const = 0;
  // The filter a toJson text describes, of the class its "type" names
  // (BloomFilterBase::fromJson, BloomFilterBase.cpp:25-43).  Its bit array is
  // zero until saveBitmapRecordBatches fills it: the bits travel separately.
  static Expected<std::shared_ptr<BloomFilterBase>> fromJson(const std::string& text);
  static Expected<std::shared_ptr<BloomFilterBase>> fromJson(const JsonValue& jObj);
  // The filter's bit array as record batches, and back
  // (BloomFilterBase.h:32-33; ArrowBloomFilter.cpp:77-116, BloomFilter.cpp:123-144).
  virtual Expected<std::vector<BitmapRecordBatch>> makeBitmapRecordBatches() const = 0;
  virtual Expected<void> saveBitmapRecordBatches(const std::vector<BitmapRecordBatch>& batches) = 0;

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
  // ArrowBloomFilter::fromJson (ArrowBloomFilter.cpp:127-149)
  static Expected<std::shared_ptr<ArrowBloomFilter>> fromJson(const JsonValue& jObj);
  // Two uint8 batches: the 136 mask bytes, then the blocks.
  Expected<std::vector<BitmapRecordBatch>> makeBitmapRecordBatches() const override;
  // Takes the blocks of a received filter (its size may differ from this
  // filter's); masks other than BloomFilterMasks() are an error.
  Expected<void> saveBitmapRecordBatches(const std::vector<BitmapRecordBatch>& batches) override;

 private:
  std::vector<std::string> columnNames_;
  int logNumBlocks_;
  std::shared_ptr<DeviceBuffer> blocks_;
};

class BloomFilter : public BloomFilterBase {
 public:
  BloomFilter(int64_t capacity, double falsePositiveRate);
  static std::shared_ptr<BloomFilter> make(int64_t capacity, double falsePositiveRate);
  // A filter with the given parameters and a zeroed bit array
  // (BloomFilter::make, BloomFilter.cpp:32-53).
  static std::shared_ptr<BloomFilter> make(int64_t capacity, double falsePositiveRate, bool valid,
                                           int64_t numHashFunctions, int64_t numBits,
                                           std::vector<fpdb_universal_hash> hashFunctions);
  // BloomFilter::fromJson (BloomFilter.cpp:166-205)
  static Expected<std::shared_ptr<BloomFilter>> fromJson(const JsonValue& jObj);

  // k, m, hash functions and a zeroed bit array (BloomFilter::init, BloomFilter.cpp:55-60).
  void init(uint32_t seed);
  Expected<void> build(const std::shared_ptr<TupleSet>& tupleSet, const std::vector<std::string>& columnNames);
  // OR of two filters with the same capacity and rate (BloomFilter::merge, BloomFilter.cpp:101-121).
  Expected<void> merge(const std::shared_ptr<BloomFilter>& other);

  double falsePositiveRate() const { return fpr_; }
  // k; an empty input keeps k but has no hash functions (BloomFilter.cpp:209-213)
  int64_t numHashFunctions() const { return numHashFunctions_; }
  int64_t numBits() const { return numBits_; }
  const std::vector<fpdb_universal_hash>& hashFunctions() const { return hashFunctions_; }
  const int64_t* deviceWords() const { return static_cast<const int64_t*>(words_->data()); }
  std::vector<int64_t> getBitArray() const;
  std::string toJson() const override;
  // One int64 batch of the bit array words.
  Expected<std::vector<BitmapRecordBatch>> makeBitmapRecordBatches() const override;
  Expected<void> saveBitmapRecordBatches(const std::vector<BitmapRecordBatch>& batches) override;

 private:
  double fpr_;
  int64_t numHashFunctions_ = 0;
  int64_t numBits_ = 0;
  std::vector<fpdb_universal_hash> hashFunctions_;
  std::shared_ptr<DeviceBuffer> words_;
};

// A filter's selection over all its input batches as int64 words in HBM,
// the bitmap of FilterPOp's bitmap pushdown (FPDBStoreFilterBitmapWrapper::bitmap_,
// fpdb-executor/src/physical/filter/FilterPOp.cpp:264-297).
class FilterBitmap {
 public:
  // ORs the n-row selection bitmap of one batch in at rowOffset (bufferBitMap).
  Expected<void> bufferBitMap(const uint8_t* d_bitmap, int64_t rowOffset, int64_t n);
  // Rows [startRowOffset, startRowOffset + numRows) as a device selection
  // bitmap of (numRows + 7) / 8 bytes (makeSelectionVector).
  Expected<std::shared_ptr<DeviceBuffer>> makeSelectionBitmap(int64_t startRowOffset, int64_t numRows) const;
  int64_t numRows() const { return numRows_; }
  std::vector<int64_t> toHost() const;
  BitmapRecordBatch toRecordBatch() const { return bitmapToRecordBatch(toHost()); }
  static Expected<FilterBitmap> fromRecordBatch(const BitmapRecordBatch& batch);

 private:
  Expected<void> reserve(int64_t words);
  std::shared_ptr<DeviceBuffer> words_;
  int64_t numWords_ = 0, capacityWords_ = 0, numRows_ = 0;
};

// Resolves key column names to columns of a TupleSet.
Expected<std::vector<fpdb_column>> keyColumns(const std::shared_ptr<TupleSet>& ts,
                                              const std::vector<std::string>& names);

}  // namespace fpdb::amd
