#include "BloomFilter.h"

#include <cstdio>
#include <cstring>
#include <sstream>

#include "Globals.h"

#include <algorithm>

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

static std::string jsonString(const std::string& v) {
  std::string o = "\"";
  for (char c : v) {
    if (c == '"' || c == '\\') o += '\\';
    o += c;
  }
  return o + "\"";
}

// ArrowBloomFilter::toJson (ArrowBloomFilter.cpp:118-125); the blocked
// filter's own object carries log_num_blocks (arrow BlockedBloomFilter::toJson,
// fpdb-tuple/src/arrow/exec/BloomFilter.cpp:289-293).
std::string ArrowBloomFilter::toJson() const {
  std::ostringstream o;
  o << "{\"type\":" << (int)type_ << ",\"capacity\":" << capacity_ << ",\"columnNames\":[";
  for (size_t i = 0; i < columnNames_.size(); ++i) o << (i ? "," : "") << jsonString(columnNames_[i]);
  o << "],\"blockedBloomFilter\":{\"log_num_blocks\":" << logNumBlocks_ << "}}";
  return o.str();
}

Expected<std::shared_ptr<ArrowBloomFilter>> ArrowBloomFilter::fromJson(const JsonValue& jObj) {
  auto capacity = jsonInt(jObj, "capacity", "ArrowBloomFilter");
  if (!capacity) return make_unexpected(capacity.error());
  if (!jObj.contains("columnNames") || jObj["columnNames"].kind != JsonValue::ARRAY)
    return make_unexpected("ColumnNames not specified in ArrowBloomFilter JSON");
  std::vector<std::string> names;
  for (auto& v : jObj["columnNames"].items) {
    if (v.kind != JsonValue::STRING) return make_unexpected("ColumnNames in ArrowBloomFilter JSON must be strings");
    names.push_back(v.s);
  }
  if (!jObj.contains("blockedBloomFilter"))
    return make_unexpected("BlockedBloomFilter not specified in ArrowBloomFilter JSON");
  auto lnb = jsonInt(jObj["blockedBloomFilter"], "log_num_blocks", "BlockedBloomFilter");
  if (!lnb) return make_unexpected(lnb.error());
  if (*lnb < 0 || *lnb > 40) return make_unexpected("log_num_blocks out of range in BlockedBloomFilter JSON");
  auto bf = make(*capacity, names);
  if (bf->valid_ && *lnb != bf->logNumBlocks_) {
    fpdb_context* ctx = defaultContext();
    bf->blocks_ = std::make_shared<DeviceBuffer>(ctx, (uint64_t)8 << *lnb);
    checkStatus(fpdb_fill(ctx, bf->blocks_->data(), 0, bf->blocks_->size()), "fpdb_fill");
  }
  bf->logNumBlocks_ = (int)*lnb;
  return bf;
}

Expected<std::vector<BitmapRecordBatch>> ArrowBloomFilter::makeBitmapRecordBatches() const {
  if (!valid_) return make_unexpected("ArrowBloomFilter: an invalid filter has no bit array");
  BitmapRecordBatch masks, blocks;
  masks.length = 136;
  blocks.length = (int64_t)8 << logNumBlocks_;
  masks.bytes.resize(136);
  blocks.bytes.resize((size_t)blocks.length);
  if (fpdb_bloom_blocked_export(defaultContext(), deviceBlocks(), logNumBlocks_, masks.bytes.data(),
                                blocks.bytes.data()) != FPDB_OK)
    return make_unexpected(fpdb_last_error());
  return std::vector<BitmapRecordBatch>{masks, blocks};
}

// ArrowBloomFilter::saveBitmapRecordBatches (ArrowBloomFilter.cpp:77-95)
Expected<void> ArrowBloomFilter::saveBitmapRecordBatches(const std::vector<BitmapRecordBatch>& batches) {
  if (batches.size() != 2)
    return make_unexpected("RecordBatch stream for ArrowBloomFilter's bitmap should contain two recordBatches");
  const auto& m = batches[0];
  const auto& b = batches[1];
  if (m.type != FPDB_TYPE_UINT8 || b.type != FPDB_TYPE_UINT8 || m.bytes.size() != 136)
    return make_unexpected("ArrowBloomFilter bitmap batches must be uint8, the first of 136 mask bytes");
  fpdb_context* ctx = defaultContext();
  int lnb = 0;
  if (fpdb_bloom_blocked_import(ctx, m.bytes.data(), b.bytes.data(), b.bytes.size(), &lnb, nullptr) != FPDB_OK)
    return make_unexpected(fpdb_last_error());
  if (!blocks_ || lnb != logNumBlocks_) blocks_ = std::make_shared<DeviceBuffer>(ctx, b.bytes.size());
  if (fpdb_bloom_blocked_import(ctx, m.bytes.data(), b.bytes.data(), b.bytes.size(), &lnb,
                                static_cast<uint64_t*>(blocks_->data())) != FPDB_OK)
    return make_unexpected(fpdb_last_error());
  logNumBlocks_ = lnb;
  valid_ = true;
  return {};
}

// ---- wire form of bitmaps ------------------------------------------------------

BitmapRecordBatch bitmapToRecordBatch(const std::vector<int64_t>& bitmap) {
  BitmapRecordBatch b;
  b.type = FPDB_TYPE_INT64;
  b.length = (int64_t)bitmap.size();
  b.bytes.resize(bitmap.size() * 8);
  if (!bitmap.empty()) std::memcpy(b.bytes.data(), bitmap.data(), b.bytes.size());
  return b;
}

Expected<std::vector<int64_t>> recordBatchToBitmap(const BitmapRecordBatch& batch) {
  if (batch.fieldName != "bitmap") return make_unexpected("Bitmap array not found in the recordBatch");
  if (batch.type != FPDB_TYPE_INT64) return make_unexpected("Type of bitmap array is not int64");
  std::vector<int64_t> out((size_t)batch.length);
  if (batch.bytes.size() != out.size() * 8) return make_unexpected("Bitmap batch has the wrong number of bytes");
  if (!out.empty()) std::memcpy(out.data(), batch.bytes.data(), batch.bytes.size());
  return out;
}

Expected<void> FilterBitmap::reserve(int64_t words) {
  if (words <= capacityWords_) return {};
  fpdb_context* ctx = defaultContext();
  const int64_t cap = std::max<int64_t>(words, capacityWords_ * 2);
  auto grown = std::make_shared<DeviceBuffer>(ctx, (uint64_t)cap * 8);
  if (fpdb_fill(ctx, grown->data(), 0, grown->size()) != FPDB_OK) return make_unexpected(fpdb_last_error());
  if (numWords_ && fpdb_copy_device(ctx, grown->data(), words_->data(), (uint64_t)numWords_ * 8) != FPDB_OK)
    return make_unexpected(fpdb_last_error());
  words_ = grown;
  capacityWords_ = cap;
  return {};
}

Expected<void> FilterBitmap::bufferBitMap(const uint8_t* d_bitmap, int64_t rowOffset, int64_t n) {
  const int64_t words = (rowOffset + n + 63) / 64;   // resize(ceil((rowOffset + n) / 64)), FilterPOp.cpp:274
  if (auto r = reserve(words); !r) return r;
  numWords_ = std::max(numWords_, words);
  numRows_ = std::max(numRows_, rowOffset + n);
  if (fpdb_bitmap_to_words(defaultContext(), d_bitmap, n, rowOffset, static_cast<int64_t*>(words_->data())) != FPDB_OK)
    return make_unexpected(fpdb_last_error());
  return {};
}

Expected<std::shared_ptr<DeviceBuffer>> FilterBitmap::makeSelectionBitmap(int64_t startRowOffset,
                                                                          int64_t numRows) const {
  fpdb_context* ctx = defaultContext();
  auto bm = std::make_shared<DeviceBuffer>(ctx, (uint64_t)(numRows + 63) / 64 * 8);
  if (numRows == 0) return bm;
  if (!words_) return make_unexpected("Bitmap not set");
  if (fpdb_bitmap_from_words(ctx, static_cast<const int64_t*>(words_->data()), numWords_, startRowOffset, numRows,
                             static_cast<uint8_t*>(bm->data())) != FPDB_OK)
    return make_unexpected(fpdb_last_error());
  return bm;
}

std::vector<int64_t> FilterBitmap::toHost() const {
  std::vector<int64_t> out((size_t)numWords_);
  if (numWords_)
    checkStatus(fpdb_copy_to_host(defaultContext(), out.data(), words_->data(), out.size() * 8), "copy");
  return out;
}

Expected<FilterBitmap> FilterBitmap::fromRecordBatch(const BitmapRecordBatch& batch) {
  auto words = recordBatchToBitmap(batch);
  if (!words) return make_unexpected(words.error());
  FilterBitmap fb;
  if (auto r = fb.reserve((int64_t)words->size()); !r) return make_unexpected(r.error());
  fb.numWords_ = (int64_t)words->size();
  fb.numRows_ = fb.numWords_ * 64;
  if (!words->empty())
    checkStatus(fpdb_copy_to_device(defaultContext(), fb.words_->data(), words->data(), words->size() * 8), "copy");
  return fb;
}

// ---- legacy filter -------------------------------------------------------------

BloomFilter::BloomFilter(int64_t capacity, double falsePositiveRate)
    : BloomFilterBase(BLOOM_FILTER, capacity, capacity <= BLOOM_FILTER_MAX_INPUT_SIZE), fpr_(falsePositiveRate) {}

std::shared_ptr<BloomFilter> BloomFilter::make(int64_t capacity, double falsePositiveRate) {
  return std::make_shared<BloomFilter>(capacity, falsePositiveRate);
}

std::shared_ptr<BloomFilter> BloomFilter::make(int64_t capacity, double falsePositiveRate, bool valid,
                                               int64_t numHashFunctions, int64_t numBits,
                                               std::vector<fpdb_universal_hash> hashFunctions) {
  auto bf = std::make_shared<BloomFilter>(capacity, falsePositiveRate);
  bf->valid_ = valid;
  bf->numHashFunctions_ = numHashFunctions;
  bf->numBits_ = numBits;
  bf->hashFunctions_ = std::move(hashFunctions);
  fpdb_context* ctx = defaultContext();
  bf->words_ = std::make_shared<DeviceBuffer>(ctx, (uint64_t)(numBits / 64 + 1) * 8);
  checkStatus(fpdb_fill(ctx, bf->words_->data(), 0, bf->words_->size()), "fpdb_fill");
  return bf;
}

Expected<std::shared_ptr<BloomFilter>> BloomFilter::fromJson(const JsonValue& jObj) {
  const std::string what = "bloom filter";
  auto capacity = jsonInt(jObj, "capacity", what);
  if (!capacity) return make_unexpected(capacity.error());
  auto fpr = jsonDouble(jObj, "falsePositiveRate", what);
  if (!fpr) return make_unexpected(fpr.error());
  auto valid = jsonBool(jObj, "valid", what);
  if (!valid) return make_unexpected(valid.error());
  auto k = jsonInt(jObj, "numHashFunctions", what);
  if (!k) return make_unexpected(k.error());
  auto numBits = jsonInt(jObj, "numBits", what);
  if (!numBits) return make_unexpected(numBits.error());
  if (*numBits < 0 || *numBits > ((int64_t)1 << 40)) return make_unexpected("numBits out of range in bloom filter JSON");
  if (!jObj.contains("hashFunctions") || jObj["hashFunctions"].kind != JsonValue::ARRAY)
    return make_unexpected("HashFunctions not specified in bloom filter JSON");
  std::vector<fpdb_universal_hash> hfs;
  for (auto& h : jObj["hashFunctions"].items) {
    // UniversalHashFunction::fromJson (UniversalHashFunction.cpp:68-92)
    fpdb_universal_hash u{};
    int64_t* dst[4] = {&u.a, &u.b, &u.m, &u.p};
    const char* keys[4] = {"a", "b", "m", "p"};
    for (int j = 0; j < 4; ++j) {
      auto v = jsonInt(h, keys[j], "universal hash function");
      if (!v) return make_unexpected(v.error());
      *dst[j] = *v;
    }
    if (u.p <= 0 || u.m <= 0 || u.m > *numBits)
      return make_unexpected("universal hash function needs p > 0 and 0 < m <= numBits");
    hfs.push_back(u);
  }
  return make(*capacity, *fpr, *valid, *k, *numBits, std::move(hfs));
}

Expected<std::shared_ptr<BloomFilterBase>> BloomFilterBase::fromJson(const JsonValue& jObj) {
  if (!jObj.contains("type")) return make_unexpected("Type not specified in BloomFilterBase JSON");
  auto type = jsonInt(jObj, "type", "BloomFilterBase");
  if (!type) return make_unexpected(type.error());
  switch (*type) {
    case BLOOM_FILTER: {
      auto r = BloomFilter::fromJson(jObj);
      if (!r) return make_unexpected(r.error());
      return std::shared_ptr<BloomFilterBase>(*r);
    }
    case ARROW_BLOOM_FILTER: {
      auto r = ArrowBloomFilter::fromJson(jObj);
      if (!r) return make_unexpected(r.error());
      return std::shared_ptr<BloomFilterBase>(*r);
    }
    default: return make_unexpected("Unknown bloom filter type: " + std::to_string(*type));
  }
}

Expected<std::shared_ptr<BloomFilterBase>> BloomFilterBase::fromJson(const std::string& text) {
  auto j = parseJson(text);
  if (!j) return make_unexpected(j.error());
  return fromJson(*j);
}

void BloomFilter::init(uint32_t seed) {
  int64_t k = 0, m = 0;
  checkStatus(fpdb_bloom_legacy_params(capacity_, fpr_, &k, &m), "fpdb_bloom_legacy_params");
  numBits_ = m;
  numHashFunctions_ = k;
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

// BloomFilter::makeBitmapRecordBatches / saveBitmapRecordBatches (BloomFilter.cpp:123-144)
Expected<std::vector<BitmapRecordBatch>> BloomFilter::makeBitmapRecordBatches() const {
  return std::vector<BitmapRecordBatch>{bitmapToRecordBatch(getBitArray())};
}

Expected<void> BloomFilter::saveBitmapRecordBatches(const std::vector<BitmapRecordBatch>& batches) {
  if (batches.size() != 1)
    return make_unexpected("RecordBatch stream for BloomFilter's bitmap should only contain one recordBatch");
  auto words = recordBatchToBitmap(batches[0]);
  if (!words) return make_unexpected(words.error());
  if ((int64_t)words->size() != numBits_ / 64 + 1)
    return make_unexpected("BloomFilter bitmap has " + std::to_string(words->size()) + " words, expected " +
                           std::to_string(numBits_ / 64 + 1));
  checkStatus(fpdb_copy_to_device(defaultContext(), words_->data(), words->data(), words->size() * 8), "copy");
  return {};
}

std::string BloomFilter::toJson() const {
  std::ostringstream o;
  char fpr[32];
  std::snprintf(fpr, sizeof fpr, "%.17g", fpr_);
  o << "{\"type\":" << (int)type_ << ",\"capacity\":" << capacity_ << ",\"falsePositiveRate\":" << fpr
    << ",\"valid\":" << (valid_ ? "true" : "false") << ",\"numHashFunctions\":" << numHashFunctions_
    << ",\"numBits\":" << numBits_ << ",\"hashFunctions\":[";
  for (size_t i = 0; i < hashFunctions_.size(); ++i) {
    const auto& h = hashFunctions_[i];
    o << (i ? "," : "") << "{\"a\":" << h.a << ",\"b\":" << h.b << ",\"m\":" << h.m << ",\"p\":" << h.p << "}";
  }
  o << "]}";
  return o.str();
}

}  // namespace fpdb::amd
