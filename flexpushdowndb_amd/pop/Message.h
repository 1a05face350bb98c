// Messages between physical operators (fpdb-executor/include/fpdb/executor/message/
// MessageType.h:10-35, TupleSetMessage.h:19, BloomFilterMessage.h:18,
// CompleteMessage, ErrorMessage, TupleSetIndexMessage.h:17), delivered in Envelopes.
#pragma once

#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "BloomFilter.h"
#include "TupleSet.h"
#include "TupleSetIndex.h"

namespace fpdb::amd {

enum class MessageType {
  START, TUPLESET, TUPLESET_INDEX, BLOOM_FILTER, COMPLETE, ERROR,
  // segment cache protocol (fpdb-executor/include/fpdb/executor/message/cache/*, ScanMessage.h)
  SCAN, LOAD_REQUEST, LOAD_RESPONSE, STORE_REQUEST, CACHE_METRICS
};

// Cache key of one column segment of a table partition
// (fpdb-cache/include/fpdb/cache/SegmentKey.h, SegmentRange.h): equality and
// hashing use the partition and the column, not the range
// (fpdb-cache/src/SegmentKey.cpp:58-72).
struct SegmentKey {
  std::string partition;
  std::string columnName;
  int64_t startOffset = 0, finishOffset = 0;
  uint64_t size = 0;    // SegmentMetadata::size: Arrow buffer bytes
  std::string toString() const {
    return "{ partition: " + partition + ", column: " + columnName + ", range: (" + std::to_string(startOffset) + ":" +
           std::to_string(finishOffset) + ") }";
  }
};

class Message {
 public:
  Message(MessageType type, std::string sender) : type_(type), sender_(std::move(sender)) {}
  virtual ~Message() = default;
  MessageType type() const { return type_; }
  const std::string& sender() const { return sender_; }
  std::string getTypeString() const;

 private:
  MessageType type_;
  std::string sender_;
};

class StartMessage : public Message {
 public:
  explicit StartMessage(std::string sender) : Message(MessageType::START, std::move(sender)) {}
};

class TupleSetMessage : public Message {
 public:
  TupleSetMessage(std::shared_ptr<TupleSet> tuples, std::string sender)
      : Message(MessageType::TUPLESET, std::move(sender)), tuples_(std::move(tuples)) {}
  const std::shared_ptr<TupleSet>& tuples() const { return tuples_; }

 private:
  std::shared_ptr<TupleSet> tuples_;
};

// A hash join build side on its way from HashJoinBuildPOp to HashJoinProbePOp
// (fpdb-executor/include/fpdb/executor/message/TupleSetIndexMessage.h:17-31).
class TupleSetIndexMessage : public Message {
 public:
  TupleSetIndexMessage(std::shared_ptr<TupleSetIndex> index, std::string sender)
      : Message(MessageType::TUPLESET_INDEX, std::move(sender)), tupleSetIndex_(std::move(index)) {}
  const std::shared_ptr<TupleSetIndex>& getTupleSetIndex() const { return tupleSetIndex_; }

 private:
  std::shared_ptr<TupleSetIndex> tupleSetIndex_;
};

class BloomFilterMessage : public Message {
 public:
  BloomFilterMessage(std::shared_ptr<BloomFilterBase> bf, std::string sender)
      : Message(MessageType::BLOOM_FILTER, std::move(sender)), bloomFilter_(std::move(bf)) {}
  const std::shared_ptr<BloomFilterBase>& getBloomFilter() const { return bloomFilter_; }

 private:
  std::shared_ptr<BloomFilterBase> bloomFilter_;
};

class CompleteMessage : public Message {
 public:
  explicit CompleteMessage(std::string sender) : Message(MessageType::COMPLETE, std::move(sender)) {}
};

class ErrorMessage : public Message {
 public:
  ErrorMessage(std::string content, std::string sender)
      : Message(MessageType::ERROR, std::move(sender)), content_(std::move(content)) {}
  const std::string& getContent() const { return content_; }

 private:
  std::string content_;
};

// ScanMessage (fpdb-executor/include/fpdb/executor/message/ScanMessage.h:17-38)
class ScanMessage : public Message {
 public:
  ScanMessage(std::vector<std::string> scanColumnNames, std::vector<std::string> projectColumnNames,
              std::string sender, bool resultNeeded)
      : Message(MessageType::SCAN, std::move(sender)), scanColumnNames_(std::move(scanColumnNames)),
        projectColumnNames_(std::move(projectColumnNames)), resultNeeded_(resultNeeded) {}
  const std::vector<std::string>& getScanColumnNames() const { return scanColumnNames_; }
  const std::vector<std::string>& getProjectColumnNames() const { return projectColumnNames_; }
  bool isResultNeeded() const { return resultNeeded_; }

 private:
  std::vector<std::string> scanColumnNames_, projectColumnNames_;
  bool resultNeeded_;
};

class LoadRequestMessage : public Message {
 public:
  LoadRequestMessage(std::vector<SegmentKey> keys, std::string sender)
      : Message(MessageType::LOAD_REQUEST, std::move(sender)), keys_(std::move(keys)) {}
  const std::vector<SegmentKey>& getSegmentKeys() const { return keys_; }

 private:
  std::vector<SegmentKey> keys_;
};

// Hit segments (device columns, pinned while held) in request order, and the
// missed keys the policy wants cached.
class LoadResponseMessage : public Message {
 public:
  LoadResponseMessage(std::vector<std::pair<SegmentKey, Column>> segments, std::vector<SegmentKey> keysToCache,
                      std::string sender)
      : Message(MessageType::LOAD_RESPONSE, std::move(sender)), segments_(std::move(segments)),
        keysToCache_(std::move(keysToCache)) {}
  const std::vector<std::pair<SegmentKey, Column>>& getSegments() const { return segments_; }
  const std::vector<SegmentKey>& getSegmentKeysToCache() const { return keysToCache_; }

 private:
  std::vector<std::pair<SegmentKey, Column>> segments_;
  std::vector<SegmentKey> keysToCache_;
};

class StoreRequestMessage : public Message {
 public:
  StoreRequestMessage(std::vector<std::pair<SegmentKey, Column>> segments, std::string sender)
      : Message(MessageType::STORE_REQUEST, std::move(sender)), segments_(std::move(segments)) {}
  const std::vector<std::pair<SegmentKey, Column>>& getSegments() const { return segments_; }

 private:
  std::vector<std::pair<SegmentKey, Column>> segments_;
};

class CacheMetricsMessage : public Message {
 public:
  CacheMetricsMessage(size_t hitNum, size_t missNum, size_t shardHitNum, size_t shardMissNum, std::string sender)
      : Message(MessageType::CACHE_METRICS, std::move(sender)), hitNum_(hitNum), missNum_(missNum),
        shardHitNum_(shardHitNum), shardMissNum_(shardMissNum) {}
  size_t getHitNum() const { return hitNum_; }
  size_t getMissNum() const { return missNum_; }
  size_t getShardHitNum() const { return shardHitNum_; }
  size_t getShardMissNum() const { return shardMissNum_; }

 private:
  size_t hitNum_, missNum_, shardHitNum_, shardMissNum_;
};

class Envelope {
 public:
  explicit Envelope(std::shared_ptr<Message> m) : message_(std::move(m)) {}
  const Message& message() const { return *message_; }
  const std::shared_ptr<Message>& ptr() const { return message_; }

 private:
  std::shared_ptr<Message> message_;
};

}  // namespace fpdb::amd
