// Segment caching around the hot path: the device-resident SegmentCache
// (fpdb-cache/include/fpdb/cache/SegmentCache.h) over fpdb_cache_* of the C
// ABI, the actor that serves it (SegmentCacheActor,
// fpdb-executor/src/cache/SegmentCacheActor.cpp), CacheLoadPOp
// (fpdb-executor/src/physical/cache/CacheLoadPOp.cpp, caching-only mode), a
// scan of host-resident columns standing in for the store's file scan
// (FileScanAbstractPOp, fpdb-executor/src/physical/file/FileScanAbstractPOp.cpp)
// and MergePOp (fpdb-executor/src/physical/merge/MergePOp.cpp).
#pragma once

#include <deque>
#include <map>
#include <optional>
#include <set>

#include "PhysicalOp.h"

namespace fpdb::amd {

// Made with make(): a loaded column keeps its cache alive (and its segment
// pinned) for as long as the column lives.
class SegmentCache : public std::enable_shared_from_this<SegmentCache> {
 public:
  // LRU replacement over `maxSize` bytes of HBM (LRUCachingPolicy).
  explicit SegmentCache(uint64_t maxSize);
  ~SegmentCache();
  SegmentCache(const SegmentCache&) = delete;
  SegmentCache& operator=(const SegmentCache&) = delete;
  static std::shared_ptr<SegmentCache> make(uint64_t maxSize) { return std::make_shared<SegmentCache>(maxSize); }

  // Copies the column into the cache (SegmentCache.cpp:21-30).
  Expected<void> store(const SegmentKey& key, const Column& column);
  // The cached column, held (pinned) for as long as the returned Column lives.
  std::optional<Column> load(const SegmentKey& key);
  unsigned long remove(const std::string& partition, const std::optional<std::string>& column = std::nullopt);
  // Missed keys the policy wants cached; LRU wants all of them
  // (LRUCachingPolicy::onToCache, LRUCachingPolicy.cpp:64-70).
  std::vector<SegmentKey> toCache(std::vector<SegmentKey> missKeys) const { return missKeys; }
  void newQuery() { clearCrtQueryMetrics(); }
  void clear();

  size_t getSize() const;
  uint64_t usedBytes() const;
  std::string showCurrentLayout() const;
  size_t hitNum() const { return hitNum_; }
  size_t missNum() const { return missNum_; }
  size_t shardHitNum() const { return shardHitNum_; }
  size_t shardMissNum() const { return shardMissNum_; }
  size_t crtQueryHitNum() const { return crtQueryHitNum_; }
  size_t crtQueryMissNum() const { return crtQueryMissNum_; }
  size_t crtQueryShardHitNum() const { return crtQueryShardHitNum_; }
  size_t crtQueryShardMissNum() const { return crtQueryShardMissNum_; }
  // SegmentCacheActor::metrics (SegmentCacheActor.cpp:63-72)
  void addMetrics(size_t hit, size_t miss, size_t shardHit, size_t shardMiss);
  void clearMetrics();
  void clearCrtQueryMetrics();

 private:
  fpdb_segment_cache* cache_ = nullptr;
  size_t hitNum_ = 0, missNum_ = 0, shardHitNum_ = 0, shardMissNum_ = 0;
  size_t crtQueryHitNum_ = 0, crtQueryMissNum_ = 0, crtQueryShardHitNum_ = 0, crtQueryShardMissNum_ = 0;
};

// The actor every CacheLoadPOp and caching scan talks to, by the name
// "SegmentCache".  It completes on START (it is a service, not a producer)
// and keeps answering afterwards.
class SegmentCacheActor : public PhysicalOp {
 public:
  explicit SegmentCacheActor(std::shared_ptr<SegmentCache> cache)
      : PhysicalOp("SegmentCache", POpType::SEGMENT_CACHE, {}, 0), cache_(std::move(cache)) {}
  std::string getTypeString() const override { return "SegmentCacheActor"; }
  void onReceive(const Envelope& msg) override;
  void clear() override {}
  const std::shared_ptr<SegmentCache>& cache() const { return cache_; }

 private:
  std::shared_ptr<SegmentCache> cache_;
};

// Caching-only CacheLoadPOp: asks the cache for every column of its segment;
// the hit columns go to the hit operator as one TupleSet, the missed ones are
// requested from the miss operator with a ScanMessage (CacheLoadPOp.cpp:115-259;
// hybrid pushdown needs the store and is not part of this layer).
class CacheLoadPOp : public PhysicalOp {
 public:
  CacheLoadPOp(const std::string& name, const std::vector<std::string>& projectColumnNames, int nodeId,
               std::vector<std::string> predicateColumnNames, std::vector<std::string> allColumnNames,
               std::string partition, int64_t startOffset, int64_t finishOffset);
  std::string getTypeString() const override { return "CacheLoadPOp"; }
  void onReceive(const Envelope& msg) override;
  void clear() override {}
  void setHitOperator(const std::shared_ptr<PhysicalOp>& op);
  void setMissOperatorToCache(const std::shared_ptr<PhysicalOp>& op);
  const std::string& getPartition() const { return partition_; }

 private:
  void onStart();
  void onCacheLoadResponse(const LoadResponseMessage& msg);

  std::vector<std::string> predicateColumnNames_, allColumnNames_;
  std::string partition_;
  int64_t startOffset_, finishOffset_;
  std::optional<std::string> hitOperatorName_, missOperatorToCacheName_;
};

// Scan of one partition whose columns sit in host memory (the store side):
// on a ScanMessage it copies the asked-for columns to the device, sends them
// on, and with toCache asks the cache to keep them
// (FileScanAbstractPOp.cpp:35-165).  scanOnStart scans every column at START.
class HostScanPOp : public PhysicalOp {
 public:
  HostScanPOp(const std::string& name, const std::vector<std::string>& projectColumnNames, int nodeId,
              Schema schema, std::vector<HostColumn> columns, std::string partition, int64_t startOffset,
              int64_t finishOffset, bool scanOnStart, bool toCache);
  std::string getTypeString() const override { return "HostScanPOp"; }
  void onReceive(const Envelope& msg) override;
  void clear() override {}
  // Bytes copied host -> device by this operator.
  uint64_t bytesRead() const { return bytesRead_; }

 private:
  Expected<std::shared_ptr<TupleSet>> readTuples(const std::vector<std::string>& columnNames);
  void readAndSendTuples(const std::vector<std::string>& columnNames);

  Schema schema_;
  std::vector<HostColumn> columns_;
  std::string partition_;
  int64_t startOffset_, finishOffset_;
  bool scanOnStart_, toCache_;
  uint64_t bytesRead_ = 0;
};

// Joins the columns of the left and right producers' TupleSets, pairwise in
// arrival order (MergePOp.cpp, MergeKernel.cpp: equal row counts, a column
// name already on the left is dropped from the right, an empty side passes
// the other through), then projects to projectColumnNames.
class MergePOp : public PhysicalOp {
 public:
  MergePOp(const std::string& name, const std::vector<std::string>& projectColumnNames, int nodeId)
      : PhysicalOp(name, POpType::MERGE, projectColumnNames, nodeId) {}
  std::string getTypeString() const override { return "MergePOp"; }
  void onReceive(const Envelope& msg) override;
  void clear() override {
    left_.clear();
    right_.clear();
  }
  void setLeftProducer(const std::shared_ptr<PhysicalOp>& op);
  void setRightProducer(const std::shared_ptr<PhysicalOp>& op);
  static Expected<std::shared_ptr<TupleSet>> merge(const std::shared_ptr<TupleSet>& a,
                                                   const std::shared_ptr<TupleSet>& b);

 private:
  void mergeReady();
  std::string leftName_, rightName_;
  std::deque<std::shared_ptr<TupleSet>> left_, right_;
};

}  // namespace fpdb::amd
