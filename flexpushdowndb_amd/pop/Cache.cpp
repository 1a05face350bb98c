#include "Cache.h"

#include <stdexcept>

namespace fpdb::amd {

// ---- SegmentCache -----------------------------------------------------------------

SegmentCache::SegmentCache(uint64_t maxSize) {
  checkStatus(fpdb_cache_create(defaultContext(), maxSize, FPDB_CACHE_LRU, &cache_), "fpdb_cache_create");
}

SegmentCache::~SegmentCache() { fpdb_cache_destroy(cache_); }

Expected<void> SegmentCache::store(const SegmentKey& key, const Column& column) {
  fpdb_column c = column.c();
  int stored = 0;
  if (fpdb_cache_store(cache_, key.partition.c_str(), key.columnName.c_str(), key.startOffset, key.finishOffset, &c,
                       &stored) != FPDB_OK)
    return make_unexpected(fpdb_last_error());
  return {};
}

std::optional<Column> SegmentCache::load(const SegmentKey& key) {
  fpdb_column c;
  int hit = 0;
  checkStatus(fpdb_cache_load(cache_, key.partition.c_str(), key.columnName.c_str(), key.startOffset,
                              key.finishOffset, &c, &hit),
              "fpdb_cache_load");
  if (!hit) return std::nullopt;
  fpdb_context* ctx = defaultContext();
  // one pin for the segment, released when both buffers are gone; the holder
  // keeps the cache itself alive until then
  auto self = shared_from_this();
  auto pinHolder = std::shared_ptr<void>(c.values, [self](void* p) { fpdb_cache_release(self->cache_, p); });
  Column col;
  col.type = (fpdb_type)c.type;
  col.length = c.length;
  col.values = std::make_shared<DeviceBuffer>(ctx, c.values, (uint64_t)c.length * typeWidth(col.type),
                                              [pinHolder] {});
  if (c.validity)
    col.validity = std::make_shared<DeviceBuffer>(ctx, c.validity, (uint64_t)(c.length + 7) / 8, [pinHolder] {});
  return col;
}

unsigned long SegmentCache::remove(const std::string& partition, const std::optional<std::string>& column) {
  int64_t removed = 0;
  checkStatus(fpdb_cache_remove(cache_, partition.c_str(), column ? column->c_str() : nullptr, &removed),
              "fpdb_cache_remove");
  return (unsigned long)removed;
}

void SegmentCache::clear() {
  checkStatus(fpdb_cache_clear(cache_), "fpdb_cache_clear");
  clearMetrics();
}

size_t SegmentCache::getSize() const {
  int64_t n = 0;
  checkStatus(fpdb_cache_info(cache_, &n, nullptr, nullptr, nullptr, nullptr), "fpdb_cache_info");
  return (size_t)n;
}

uint64_t SegmentCache::usedBytes() const {
  uint64_t used = 0;
  checkStatus(fpdb_cache_info(cache_, nullptr, &used, nullptr, nullptr, nullptr), "fpdb_cache_info");
  return used;
}

std::string SegmentCache::showCurrentLayout() const {
  uint64_t need = 0;
  checkStatus(fpdb_cache_layout(cache_, nullptr, 0, &need), "fpdb_cache_layout");
  std::string s(need, '\0');
  checkStatus(fpdb_cache_layout(cache_, s.data(), need, &need), "fpdb_cache_layout");
  s.resize(need ? need - 1 : 0);
  return s;
}

void SegmentCache::addMetrics(size_t hit, size_t miss, size_t shardHit, size_t shardMiss) {
  hitNum_ += hit;
  missNum_ += miss;
  shardHitNum_ += shardHit;
  shardMissNum_ += shardMiss;
  crtQueryHitNum_ += hit;
  crtQueryMissNum_ += miss;
  crtQueryShardHitNum_ += shardHit;
  crtQueryShardMissNum_ += shardMiss;
}

void SegmentCache::clearMetrics() {
  hitNum_ = missNum_ = shardHitNum_ = shardMissNum_ = 0;
  clearCrtQueryMetrics();
}

void SegmentCache::clearCrtQueryMetrics() {
  crtQueryHitNum_ = crtQueryMissNum_ = crtQueryShardHitNum_ = crtQueryShardMissNum_ = 0;
}

// ---- SegmentCacheActor ---------------------------------------------------------------

// SegmentCacheActor::load / store / metrics (SegmentCacheActor.cpp:11-72).  In
// caching-only mode every missed key is to be loaded and cached.
void SegmentCacheActor::onReceive(const Envelope& msg) {
  const auto& m = msg.message();
  switch (m.type()) {
    case MessageType::START:
      ctx()->notifyComplete();
      break;
    case MessageType::LOAD_REQUEST: {
      const auto& req = static_cast<const LoadRequestMessage&>(m);
      std::vector<std::pair<SegmentKey, Column>> hits;
      std::vector<SegmentKey> misses;
      for (const auto& key : req.getSegmentKeys()) {
        auto col = cache_->load(key);
        if (col)
          hits.emplace_back(key, *col);
        else
          misses.push_back(key);
      }
      ctx()->tell(std::make_shared<LoadResponseMessage>(std::move(hits), cache_->toCache(std::move(misses)), name_),
                  {m.sender()});
      break;
    }
    case MessageType::STORE_REQUEST: {
      for (const auto& [key, col] : static_cast<const StoreRequestMessage&>(m).getSegments()) {
        auto r = cache_->store(key, col);
        if (!r) ctx()->notifyError(r.error());
      }
      break;
    }
    case MessageType::CACHE_METRICS: {
      const auto& cm = static_cast<const CacheMetricsMessage&>(m);
      cache_->addMetrics(cm.getHitNum(), cm.getMissNum(), cm.getShardHitNum(), cm.getShardMissNum());
      break;
    }
    default:
      break;
  }
}

// ---- CacheLoadPOp -----------------------------------------------------------------------

CacheLoadPOp::CacheLoadPOp(const std::string& name, const std::vector<std::string>& projectColumnNames, int nodeId,
                           std::vector<std::string> predicateColumnNames, std::vector<std::string> allColumnNames,
                           std::string partition, int64_t startOffset, int64_t finishOffset)
    : PhysicalOp(name, POpType::CACHE_LOAD, projectColumnNames, nodeId),
      predicateColumnNames_(std::move(predicateColumnNames)), allColumnNames_(std::move(allColumnNames)),
      partition_(std::move(partition)), startOffset_(startOffset), finishOffset_(finishOffset) {}

void CacheLoadPOp::setHitOperator(const std::shared_ptr<PhysicalOp>& op) {
  hitOperatorName_ = op->name();
  produce(op);
}

void CacheLoadPOp::setMissOperatorToCache(const std::shared_ptr<PhysicalOp>& op) {
  missOperatorToCacheName_ = op->name();
  produce(op);
}

void CacheLoadPOp::onReceive(const Envelope& msg) {
  const auto& m = msg.message();
  if (m.type() == MessageType::START)
    onStart();
  else if (m.type() == MessageType::LOAD_RESPONSE)
    onCacheLoadResponse(static_cast<const LoadResponseMessage&>(m));
  else
    ctx()->notifyError("Unrecognized message type " + m.getTypeString());
}

// CacheLoadPOp::onStart + CacheHelper::requestLoadSegmentsFromCache
// (CacheLoadPOp.cpp:75-113, CacheHelper.cpp:10-33)
void CacheLoadPOp::onStart() {
  if (!hitOperatorName_) {
    ctx()->notifyError("Hit consumer not set ");
    return;
  }
  if (!missOperatorToCacheName_) {
    ctx()->notifyError("Miss caching consumer not set");
    return;
  }
  std::vector<SegmentKey> keys;
  for (const auto& c : allColumnNames_) keys.push_back(SegmentKey{partition_, c, startOffset_, finishOffset_, 0});
  ctx()->tell(std::make_shared<LoadRequestMessage>(std::move(keys), name_), {"SegmentCache"});
}

// CacheLoadPOp::onCacheLoadResponse, caching-only branch (CacheLoadPOp.cpp:115-259)
void CacheLoadPOp::onCacheLoadResponse(const LoadResponseMessage& msg) {
  Schema hitSchema;
  std::vector<Column> hitColumns;
  for (const auto& [key, col] : msg.getSegments()) {
    hitSchema.fields.push_back(Field{key.columnName, col.type, col.validity != nullptr});
    hitColumns.push_back(col);
  }
  std::vector<std::string> missCachingColumnNames;
  for (const auto& key : msg.getSegmentKeysToCache()) missCachingColumnNames.push_back(key.columnName);

  ctx()->tell(std::make_shared<TupleSetMessage>(TupleSet::make(hitSchema, hitColumns), name_), {*hitOperatorName_});
  ctx()->tell(std::make_shared<ScanMessage>(std::vector<std::string>{}, missCachingColumnNames, name_, true),
              {*missOperatorToCacheName_});

  const size_t hitNum = hitColumns.size();
  const size_t missNum = allColumnNames_.size() - hitNum;
  const bool shardHit = hitNum == allColumnNames_.size();
  ctx()->tell(std::make_shared<CacheMetricsMessage>(hitNum, missNum, shardHit ? 1 : 0, shardHit ? 0 : 1, name_),
              {"SegmentCache"});
  ctx()->notifyComplete();
}

// ---- HostScanPOp -------------------------------------------------------------------------

HostScanPOp::HostScanPOp(const std::string& name, const std::vector<std::string>& projectColumnNames, int nodeId,
                         Schema schema, std::vector<HostColumn> columns, std::string partition, int64_t startOffset,
                         int64_t finishOffset, bool scanOnStart, bool toCache)
    : PhysicalOp(name, POpType::FILE_SCAN, projectColumnNames, nodeId), schema_(std::move(schema)),
      columns_(std::move(columns)), partition_(std::move(partition)), startOffset_(startOffset),
      finishOffset_(finishOffset), scanOnStart_(scanOnStart), toCache_(toCache) {}

void HostScanPOp::onReceive(const Envelope& msg) {
  const auto& m = msg.message();
  switch (m.type()) {
    case MessageType::START:
      if (scanOnStart_) readAndSendTuples(projectColumnNames_);
      break;
    case MessageType::SCAN: {
      const auto& sm = static_cast<const ScanMessage&>(m);
      if (sm.isResultNeeded()) {
        readAndSendTuples(sm.getProjectColumnNames());
      } else {
        // an empty TupleSet first, then the read only to cache (FileScanAbstractPOp.cpp:57-71)
        ctx()->tell(std::make_shared<TupleSetMessage>(TupleSet::make(Schema{}, {}), name_));
        ctx()->notifyComplete();
        auto r = readTuples(sm.getProjectColumnNames());
        if (!r) ctx()->notifyError(r.error());
      }
      break;
    }
    case MessageType::COMPLETE:
      break;
    default:
      ctx()->notifyError("Unrecognized message type " + m.getTypeString());
  }
}

Expected<std::shared_ptr<TupleSet>> HostScanPOp::readTuples(const std::vector<std::string>& columnNames) {
  Schema s;
  std::vector<HostColumn> hcols;
  for (const auto& name : columnNames) {
    int i = schema_.getFieldIndex(name);
    if (i < 0) return make_unexpected("Column '" + name + "' not found in partition " + partition_);
    s.fields.push_back(schema_.fields[i]);
    hcols.push_back(columns_[i]);
    bytesRead_ += columns_[i].bytes.size() + (columns_[i].valid.empty() ? 0 : (columns_[i].length + 7) / 8);
  }
  std::shared_ptr<TupleSet> ts;
  try {
    ts = TupleSet::fromHost(defaultContext(), s, hcols);
  } catch (const std::exception& e) {
    return make_unexpected(e.what());
  }
  if (toCache_ && ts->numColumns() > 0) {
    // CacheHelper::requestStoreSegmentsInCache (CacheHelper.cpp:35-61)
    std::vector<std::pair<SegmentKey, Column>> segments;
    for (int c = 0; c < ts->numColumns(); ++c) {
      const Column& col = ts->column(c);
      const uint64_t size = (uint64_t)col.length * typeWidth(col.type) + (col.validity ? (col.length + 7) / 8 : 0);
      segments.emplace_back(SegmentKey{partition_, s.fields[c].name, startOffset_, finishOffset_, size}, col);
    }
    ctx()->tell(std::make_shared<StoreRequestMessage>(std::move(segments), name_), {"SegmentCache"});
  }
  return ts;
}

void HostScanPOp::readAndSendTuples(const std::vector<std::string>& columnNames) {
  auto r = readTuples(columnNames);
  if (!r) {
    ctx()->notifyError(r.error());
    return;
  }
  ctx()->tell(std::make_shared<TupleSetMessage>(*r, name_));
  ctx()->notifyComplete();
}

// ---- MergePOp ----------------------------------------------------------------------------

void MergePOp::setLeftProducer(const std::shared_ptr<PhysicalOp>& op) {
  leftName_ = op->name();
  consume(op);
}

void MergePOp::setRightProducer(const std::shared_ptr<PhysicalOp>& op) {
  rightName_ = op->name();
  consume(op);
}

Expected<std::shared_ptr<TupleSet>> MergePOp::merge(const std::shared_ptr<TupleSet>& a,
                                                    const std::shared_ptr<TupleSet>& b) {
  if (a->numColumns() > 0 && b->numColumns() > 0 && a->numRows() != b->numRows())
    return make_unexpected("Cannot merge TupleSets, number of rows must be equal. tupleSet1.numRows: " +
                           std::to_string(a->numRows()) + " != tupleSet2.numRows: " + std::to_string(b->numRows()));
  Schema s = a->schema();
  std::vector<Column> cols = a->columns();
  for (int i = 0; i < b->numColumns(); ++i) {
    const auto& f = b->schema().fields[i];
    if (s.getFieldIndex(f.name) >= 0) continue;
    s.fields.push_back(f);
    cols.push_back(b->column(i));
  }
  return TupleSet::make(s, cols);
}

void MergePOp::mergeReady() {
  while (!left_.empty() && !right_.empty()) {
    auto r = merge(left_.front(), right_.front());
    left_.pop_front();
    right_.pop_front();
    if (!r) {
      ctx()->notifyError(r.error() + ".\n leftOp: " + leftName_ + "\n rightOp: " + rightName_);
      return;
    }
    auto projected = projectColumnNames_.empty() ? *r : (*r)->projectExist(projectColumnNames_);
    if (!projected) {
      ctx()->notifyError(projected.error());
      return;
    }
    ctx()->tell(std::make_shared<TupleSetMessage>(*projected, name_));
  }
}

void MergePOp::onReceive(const Envelope& msg) {
  const auto& m = msg.message();
  if (m.type() == MessageType::TUPLESET) {
    const auto& tm = static_cast<const TupleSetMessage&>(m);
    if (m.sender() == leftName_)
      left_.push_back(tm.tuples());
    else if (m.sender() == rightName_)
      right_.push_back(tm.tuples());
    else
      ctx()->notifyError("Unrecognized producer " + m.sender() + ", left: " + leftName_ + ", right: " + rightName_);
    mergeReady();
  } else if (m.type() == MessageType::COMPLETE) {
    if (!ctx()->isComplete() && ctx()->operatorMap().allComplete(POpRelationshipType::Producer))
      ctx()->notifyComplete();
  }
}

}  // namespace fpdb::amd
