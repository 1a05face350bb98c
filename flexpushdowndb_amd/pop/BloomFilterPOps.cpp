#include "BloomFilterPOps.h"

#include "Globals.h"

namespace fpdb::amd {

// ---- kernels -----------------------------------------------------------------

BloomFilterCreateKernel::BloomFilterCreateKernel(std::vector<std::string> columnNames,
                                                 double desiredFalsePositiveRate, bool arrowImpl)
    : columnNames_(std::move(columnNames)), fpr_(desiredFalsePositiveRate), arrowImpl_(arrowImpl) {}

Expected<void> BloomFilterCreateKernel::bufferTupleSet(const std::shared_ptr<TupleSet>& tupleSet) {
  received_.push_back(tupleSet);
  return {};
}

Expected<void> BloomFilterCreateKernel::buildBloomFilter() {
  if (received_.empty()) return make_unexpected("No tupleSet received");
  int64_t capacity = 0;
  for (auto& ts : received_) capacity += ts->numRows();
  if (arrowImpl_) {
    // BloomFilterCreateArrowKernel::buildBloomFilter (BloomFilterCreateArrowKernel.cpp:17-34)
    auto bf = ArrowBloomFilter::make(capacity, columnNames_);
    bloomFilter_ = bf;
    if (!bf->valid()) return {};
    for (auto& ts : received_) {
      auto r = bf->insert(ts);
      if (!r) return r;
    }
  } else {
    // BloomFilterCreateKernel::buildBloomFilter (BloomFilterCreateKernel.cpp:24-78)
    auto bf = BloomFilter::make(capacity, fpr_);
    bloomFilter_ = bf;
    if (!bf->valid()) return {};
    bf->init(LEGACY_BLOOM_FILTER_SEED);
    // one device build over the concatenated input (OR of per-batch builds)
    auto all = TupleSet::concatenate(received_);
    if (!all) return make_unexpected(all.error());
    auto r = bf->build(*all, columnNames_);
    if (!r) return r;
  }
  return {};
}

void BloomFilterCreateKernel::clear() {
  received_.clear();
  bloomFilter_.reset();
}

Expected<std::shared_ptr<TupleSet>> BloomFilterUseKernel::filter(const std::shared_ptr<TupleSet>& tupleSet,
                                                                 const std::shared_ptr<BloomFilterBase>& bloomFilter,
                                                                 const std::vector<std::string>& columnNames) {
  if (tupleSet->numRows() == 0 || !bloomFilter->valid()) return tupleSet;
  fpdb_context* ctx = defaultContext();
  int64_t n = tupleSet->numRows();
  std::vector<fpdb_column> in;
  std::vector<Column> outCols;
  std::vector<fpdb_column> out;
  for (int i = 0; i < tupleSet->numColumns(); ++i) {
    const Column& c = tupleSet->column(i);
    in.push_back(c.c());
    outCols.push_back(Column::allocateLike(ctx, c, n));
    out.push_back(outCols.back().c());
  }
  bool anyUtf8 = false;
  for (auto& c : tupleSet->columns()) anyUtf8 = anyUtf8 || c.isUtf8();
  std::vector<int> keyIndex;
  for (auto& name : columnNames) {
    int i = tupleSet->schema().getFieldIndex(name);
    if (i < 0) return make_unexpected("Column '" + name + "' does not exist");
    keyIndex.push_back(i);
  }
  int64_t kept = 0;
  if (bloomFilter->getType() == ARROW_BLOOM_FILTER) {
    // BloomFilterUseKernel.cpp:98-138,180-207: hash, Find, arrow::compute::Filter
    auto bf = std::static_pointer_cast<ArrowBloomFilter>(bloomFilter);
    fpdb_status s;
    if (!anyUtf8) {
      s = fpdb_bloom_blocked_filter(ctx, bf->deviceBlocks(), bf->logNumBlocks(), in.data(), (int)in.size(),
                                    keyIndex.data(), (int)keyIndex.size(), n, nullptr, out.data(), &kept);
    } else {
      // the fused pass moves fixed-width columns only: Find, then a compaction
      // that carries the utf8 columns too
      std::vector<fpdb_column> keys;
      for (int i : keyIndex) keys.push_back(in[i]);
      DeviceBuffer bitmap(ctx, (uint64_t)((n + 63) / 64) * 8);
      s = fpdb_bloom_blocked_find(ctx, bf->deviceBlocks(), bf->logNumBlocks(), keys.data(), (int)keys.size(), n,
                                  static_cast<uint8_t*>(bitmap.data()), nullptr);
      if (s == FPDB_OK)
        s = fpdb_filter_compact(ctx, in.data(), (int)in.size(), n, static_cast<uint8_t*>(bitmap.data()), out.data(),
                                &kept);
    }
    if (s != FPDB_OK) return make_unexpected(fpdb_last_error());
  } else {
    // BloomFilterUseKernel.cpp:140-178: contains() per row, selection vector, filter
    auto bf = std::static_pointer_cast<BloomFilter>(bloomFilter);
    std::vector<fpdb_column> keys;
    for (int i : keyIndex) keys.push_back(in[i]);
    DeviceBuffer bitmap(ctx, (uint64_t)((n + 63) / 64) * 8);
    fpdb_status s = fpdb_bloom_legacy_find(ctx, bf->hashFunctions().data(), (int)bf->hashFunctions().size(),
                                           bf->capacity(), bf->numBits(), bf->deviceWords(), keys.data(),
                                           (int)keys.size(), n, static_cast<uint8_t*>(bitmap.data()), nullptr);
    if (s == FPDB_OK)
      s = fpdb_filter_compact(ctx, in.data(), (int)in.size(), n, static_cast<uint8_t*>(bitmap.data()), out.data(),
                              &kept);
    if (s != FPDB_OK) return make_unexpected(fpdb_last_error());
  }
  for (auto& c : outCols) c.length = kept;
  return TupleSet::make(tupleSet->schema(), outCols);
}

// ---- create POp ----------------------------------------------------------------

BloomFilterCreatePOp::BloomFilterCreatePOp(const std::string& name, const std::vector<std::string>& projectColumnNames,
                                           int nodeId, const std::vector<std::string>& bloomFilterColumnNames,
                                           double desiredFalsePositiveRate)
    : PhysicalOp(name, POpType::BLOOM_FILTER_CREATE, projectColumnNames, nodeId),
      kernel_(bloomFilterColumnNames, desiredFalsePositiveRate, USE_ARROW_BLOOM_FILTER_IMPL) {}

void BloomFilterCreatePOp::produce(const std::shared_ptr<PhysicalOp>& op) {
  passTupleSetConsumers_.insert(op->name());
  PhysicalOp::produce(op);
}

void BloomFilterCreatePOp::addBloomFilterUsePOp(const std::shared_ptr<PhysicalOp>& bloomFilterUsePOp) {
  bloomFilterUsePOps_.insert(bloomFilterUsePOp->name());
  PhysicalOp::produce(bloomFilterUsePOp);
}

void BloomFilterCreatePOp::onReceive(const Envelope& envelope) {
  const auto& msg = envelope.message();
  if (msg.type() == MessageType::START) {
    return;
  } else if (msg.type() == MessageType::TUPLESET) {
    onTupleSet(static_cast<const TupleSetMessage&>(msg));
  } else if (msg.type() == MessageType::COMPLETE) {
    onComplete(static_cast<const CompleteMessage&>(msg));
  } else {
    ctx()->notifyError("Unrecognized message type " + msg.getTypeString());
  }
}

// BloomFilterCreatePOp.cpp:95-111: buffer, and pass the input to the other consumers
void BloomFilterCreatePOp::onTupleSet(const TupleSetMessage& msg) {
  auto tupleSet = msg.tuples();
  auto r = kernel_.bufferTupleSet(tupleSet);
  if (!r) ctx()->notifyError(r.error());
  if (!passTupleSetConsumers_.empty())
    ctx()->tell(std::make_shared<TupleSetMessage>(tupleSet, name_), passTupleSetConsumers_);
}

// BloomFilterCreatePOp.cpp:113-137: build once every producer completed
void BloomFilterCreatePOp::onComplete(const CompleteMessage&) {
  if (!ctx()->isComplete() && ctx()->operatorMap().allComplete(POpRelationshipType::Producer)) {
    auto r = kernel_.buildBloomFilter();
    if (!r) {
      ctx()->notifyError(r.error());
      return;
    }
    auto bf = kernel_.getBloomFilter();
    if (!bf.has_value()) {
      ctx()->notifyError("Bloom filter not created on complete");
      return;
    }
    ctx()->tell(std::make_shared<BloomFilterMessage>(*bf, name_), bloomFilterUsePOps_);
    ctx()->notifyComplete();
  }
}

// ---- use POp -------------------------------------------------------------------

BloomFilterUsePOp::BloomFilterUsePOp(const std::string& name, const std::vector<std::string>& projectColumnNames,
                                     int nodeId, const std::vector<std::string>& bloomFilterColumnNames)
    : PhysicalOp(name, POpType::BLOOM_FILTER_USE, projectColumnNames, nodeId),
      bloomFilterColumnNames_(bloomFilterColumnNames) {}

void BloomFilterUsePOp::onReceive(const Envelope& envelope) {
  const auto& msg = envelope.message();
  if (msg.type() == MessageType::START) {
    return;
  } else if (msg.type() == MessageType::TUPLESET) {
    onTupleSet(static_cast<const TupleSetMessage&>(msg));
  } else if (msg.type() == MessageType::BLOOM_FILTER) {
    onBloomFilter(static_cast<const BloomFilterMessage&>(msg));
  } else if (msg.type() == MessageType::COMPLETE) {
    onComplete(static_cast<const CompleteMessage&>(msg));
  } else {
    ctx()->notifyError("Unrecognized message type " + msg.getTypeString());
  }
}

// BloomFilterUsePOp.cpp:71-96: buffer until the filter arrives
void BloomFilterUsePOp::onTupleSet(const TupleSetMessage& msg) {
  auto tupleSet = msg.tuples();
  numRowsInput_ += tupleSet->numRows();
  if (!receivedTupleSet_.has_value()) {
    receivedTupleSet_ = std::make_shared<TupleSet>(tupleSet->schema(), tupleSet->columns());
  } else {
    auto r = (*receivedTupleSet_)->append(tupleSet);
    if (!r) ctx()->notifyError(r.error());
  }
  if (bloomFilter_.has_value()) {
    auto r = filterAndSend();
    if (!r) ctx()->notifyError(r.error());
  }
}

void BloomFilterUsePOp::onBloomFilter(const BloomFilterMessage& msg) {
  bloomFilter_ = msg.getBloomFilter();
  if (receivedTupleSet_.has_value()) {
    auto r = filterAndSend();
    if (!r) ctx()->notifyError(r.error());
  }
}

void BloomFilterUsePOp::onComplete(const CompleteMessage&) {
  if (!ctx()->isComplete() && ctx()->operatorMap().allComplete(POpRelationshipType::Producer)) {
    ctx()->notifyComplete();
  }
}

Expected<void> BloomFilterUsePOp::filterAndSend() {
  if (!receivedTupleSet_.has_value()) return make_unexpected("No tupleSet received");
  if (!bloomFilter_.has_value()) return make_unexpected("No bloom filter received");
  auto filtered = BloomFilterUseKernel::filter(*receivedTupleSet_, *bloomFilter_, bloomFilterColumnNames_);
  if (!filtered) return make_unexpected(filtered.error());
  numRowsOutput_ += (*filtered)->numRows();
  ctx()->tell(std::make_shared<TupleSetMessage>(*filtered, name_));
  receivedTupleSet_.reset();
  return {};
}

void BloomFilterUsePOp::clear() {
  receivedTupleSet_.reset();
  bloomFilter_.reset();
}

}  // namespace fpdb::amd
