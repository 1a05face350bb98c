#include "BloomFilter.h"

#include <sstream>

#include "Globals.h"

namespace fpdb::amd {

Expected<std::vector<fpdb_column>> keyColumns(const std::shared_ptr<TupleSet>& ts,
                                              const std::vector<std::string>& names) {
  std::vector<fpdb_column> cols;
  for (auto& n : names) {
    int i = ts->schema().getFieldIndex(n);
    if (i < 0) return make_unexpected("Column '" + n + "' does not exist");
    cols.push_back(ts->column(i).c());
  }
  return cols;
}

// ---- blocked filter ----------------------------------------------------------

ArrowBloomFilter::ArrowBloomFilter(int64_t capacity, std::vector<std::string> columnNames)
    : BloomFilterBase(ARROW_BLOOM_FILTER, capacity, capacity <= BLOOM_FILTER_MAX_INPUT_SIZE),
      columnNames_(std::move(columnNames)) {
  logNumBlocks_ = fpdb_bloom_blocked_log_num_blocks(capacity);
  fpdb_context* ctx = defaultContext();
  if (valid_) {
    blocks_ = std::make_shared<DeviceBuffer>(ctx, (uint64_t)8 << logNumBlocks_);
    checkStatus(fpdb_fill(ctx, blocks_->data(), 0, blocks_->size()), "fpdb_fill");
  }
}

std::shared_ptr<ArrowBloomFilter> ArrowBloomFilter::make(int64_t capacity,
                                                         const std::vector<std::string>& columnNames) {
  return std::make_shared<ArrowBloomFilter>(capacity, columnNames);
}

Expected<void> ArrowBloomFilter::insert(const std::shared_ptr<TupleSet>& tupleSet) {
  if (!valid_) return {};
  auto cols = keyColumns(tupleSet, columnNames_);
  if (!cols) return make_unexpected(cols.error());
  fpdb_status s = fpdb_bloom_blocked_insert(defaultContext(), cols->data(), (int)cols->size(),
                                            tupleSet->numRows(), nullptr, logNumBlocks_,
                                            static_cast<uint64_t*>(blocks_->data()));
  if (s != FPDB_OK) return make_unexpected(fpdb_last_error());
  return {};
}

Expected<void> ArrowBloomFilter::build(const std::shared_ptr<TupleSet>& tupleSet) { return insert(tupleSet); }

std::vector<uint64_t> ArrowBloomFilter::blocksToHost() const {
  std::vector<uint64_t> out;
  if (!valid_) return out;
  out.resize((size_t)1 << logNumBlocks_);
  checkStatus(fpdb_copy_to_host(defaultContext(), out.data(), blocks_->data(), out.size() * 8), "copy");
  return out;
}

std::vector<uint8_t> ArrowBloomFilter::masksBytes() const {
  std::vector<uint8_t> m(136);
  checkStatus(fpdb_bloom_masks(m.data()), "fpdb_bloom_masks");
  return m;
}

std::string ArrowBloomFilter::toJson() const {
  std::ostringstream o;
  o << "{\"type\":" << (int)type_ << ",\"capacity\":" << capacity_ << ",\"columnNames\":[";
  for (size_t i = 0; i < columnNames_.size(); ++i) o << (i ? "," : "") << "\"" << columnNames_[i] << "\"";
  o << "],\"blockedBloomFilter\":{\"logNumBlocks\":" << logNumBlocks_ << "}}";
  return o.str();
}

// ---- legacy filter -------------------------------------------------------------

BloomFilter::BloomFilter(int64_t capacity, double falsePositiveRate)
    : BloomFilterBase(BLOOM_FILTER, capacity, capacity <= BLOOM_FILTER_MAX_INPUT_SIZE), fpr_(falsePositiveRate) {}

std::shared_ptr<BloomFilter> BloomFilter::make(int64_t capacity, double falsePositiveRate) {
  return std::make_shared<BloomFilter>(capacity, falsePositiveRate);
}

void BloomFilter::init(uint32_t seed) {
  int64_t k = 0, m = 0;
  checkStatus(fpdb_bloom_legacy_params(capacity_, fpr_, &k, &m), "fpdb_bloom_legacy_params");
  numBits_ = m;
  hashFunctions_.clear();
  if (capacity_ > 0) {   // no hash functions for an empty input (BloomFilter.cpp:209-213)
    hashFunctions_.resize((size_t)k);
    checkStatus(fpdb_bloom_legacy_make_hash_functions(seed, m, (int)k, hashFunctions_.data()),
                "fpdb_bloom_legacy_make_hash_functions");
  }
  fpdb_context* ctx = defaultContext();
  words_ = std::make_shared<DeviceBuffer>(ctx, (uint64_t)(numBits_ / 64 + 1) * 8);
  checkStatus(fpdb_fill(ctx, words_->data(), 0, words_->size()), "fpdb_fill");
}

Expected<void> BloomFilter::build(const std::shared_ptr<TupleSet>& tupleSet,
                                  const std::vector<std::string>& columnNames) {
  auto cols = keyColumns(tupleSet, columnNames);
  if (!cols) return make_unexpected(cols.error());
  fpdb_status s = fpdb_bloom_legacy_build(defaultContext(), cols->data(), (int)cols->size(), tupleSet->numRows(),
                                          nullptr, hashFunctions_.data(), (int)hashFunctions_.size(), numBits_,
                                          static_cast<int64_t*>(words_->data()));
  if (s != FPDB_OK) return make_unexpected(fpdb_last_error());
  return {};
}

Expected<void> BloomFilter::merge(const std::shared_ptr<BloomFilter>& other) {
  if (capacity_ != other->capacity_)
    return make_unexpected("Capacity mismatch, " + std::to_string(capacity_) + " vs " + std::to_string(other->capacity_));
  if (fpr_ != other->fpr_)
    return make_unexpected("FalsePositiveRate mismatch");
  auto a = getBitArray();
  auto b = other->getBitArray();
  for (size_t i = 0; i < a.size() && i < b.size(); ++i) a[i] |= b[i];
  checkStatus(fpdb_copy_to_device(defaultContext(), words_->data(), a.data(), a.size() * 8), "copy");
  return {};
}

std::vector<int64_t> BloomFilter::getBitArray() const {
  std::vector<int64_t> out((size_t)(numBits_ / 64 + 1));
  checkStatus(fpdb_copy_to_host(defaultContext(), out.data(), words_->data(), out.size() * 8), "copy");
  return out;
}

std::string BloomFilter::toJson() const {
  std::ostringstream o;
  o << "{\"type\":" << (int)type_ << ",\"capacity\":" << capacity_ << ",\"falsePositiveRate\":" << fpr_
    << ",\"valid\":" << (valid_ ? "true" : "false") << ",\"numHashFunctions\":" << hashFunctions_.size()
    << ",\"numBits\":" << numBits_ << ",\"hashFunctions\":[";
  for (size_t i = 0; i < hashFunctions_.size(); ++i) {
    const auto& h = hashFunctions_[i];
    o << (i ? "," : "") << "{\"a\":" << h.a << ",\"b\":" << h.b << ",\"m\":" << h.m << ",\"p\":" << h.p << "}";
  }
  o << "]}";
  return o.str();
}

}  // namespace fpdb::amd
