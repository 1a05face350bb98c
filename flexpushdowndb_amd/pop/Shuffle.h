// Hash shuffle operator: ShufflePOp + ShuffleKernel2
// (fpdb-executor/include/fpdb/executor/physical/shuffle/ShufflePOp.h,
// ShuffleKernel2.h:31-37) over fpdb_shuffle of the C ABI.
#pragma once

#include <optional>

#include "PhysicalOp.h"

namespace fpdb::amd {

class ShuffleKernel2 {
 public:
  // numSlots TupleSets, slot p holding the rows whose key-column hash modulo
  // numSlots is p, in input order.  The slots are slices of one device
  // TupleSet (ShuffleKernel2.cpp:28-126).
  static Expected<std::vector<std::shared_ptr<TupleSet>>> shuffle(const std::vector<std::string>& columnNames,
                                                                  size_t numSlots,
                                                                  const std::shared_ptr<TupleSet>& tupleSet);
};

class ShufflePOp : public PhysicalOp {
 public:
  ShufflePOp(const std::string& name, const std::vector<std::string>& projectColumnNames, int nodeId,
             std::vector<std::string> shuffleColumnNames);
  ShufflePOp(const std::string& name, const std::vector<std::string>& projectColumnNames, int nodeId,
             std::vector<std::string> shuffleColumnNames, std::vector<std::string> consumerVec);
  std::string getTypeString() const override { return "ShufflePOp"; }
  void onReceive(const Envelope& msg) override;
  void clear() override { buffers_.clear(); }
  // A consumer added through produce() is also the next slot (ShufflePOp.cpp:70-73).
  void produce(const std::shared_ptr<PhysicalOp>& op) override;

  const std::vector<std::string>& getShuffleColumnNames() const { return shuffleColumnNames_; }
  const std::vector<std::string>& getConsumerVec() const { return consumerVec_; }
  void setConsumerVec(const std::vector<std::string>& consumerVec) { consumerVec_ = consumerVec; }
  void addToConsumerVec(const std::shared_ptr<PhysicalOp>& op) { consumerVec_.push_back(op->name()); }
  void clearConsumerVec() { consumerVec_.clear(); }

 private:
  void onStart();
  void onTupleSet(const TupleSetMessage& msg);
  void onComplete(const CompleteMessage& msg);
  Expected<void> buffer(const std::shared_ptr<TupleSet>& tupleSet, int partitionIndex);
  Expected<void> send(int partitionIndex, bool force);

  std::vector<std::string> shuffleColumnNames_;
  std::vector<std::string> consumerVec_;
  std::vector<std::optional<std::shared_ptr<TupleSet>>> buffers_;
};

}  // namespace fpdb::amd
