// The two-operator hash join: HashJoinBuildPOp buffers the build side into a
// TupleSetIndex and ships it in TupleSetIndexMessages; HashJoinProbePOp joins
// every index that arrives with the probe rows seen so far and every probe
// batch with the indexes seen so far, and adds the outer/semi rows once all
// producers are complete
// (fpdb-executor/include/fpdb/executor/physical/join/hashjoin/HashJoinBuildPOp.h:18-45,
// HashJoinProbePOp.h:19-54, HashJoinProbeAbstractKernel.h:19-63,
// HashJoinProbeKernel.h:16-52, HashSemiJoinProbeKernel.h:16-47).
// Output columns: the build side's needed columns first, then the probe
// side's (HashJoinProbeAbstractKernel.cpp:60-83).
#pragma once

#include <optional>
#include <set>

#include "HashJoin.h"

namespace fpdb::amd {

// HashJoinBuildKernel (HashJoinBuildKernel.cpp:13-51)
class HashJoinBuildKernel {
 public:
  explicit HashJoinBuildKernel(std::vector<std::string> columnNames);
  Expected<void> put(const std::shared_ptr<TupleSet>& tupleSet);
  int64_t size() const { return tupleSetIndex_ ? tupleSetIndex_->size() : 0; }
  const std::shared_ptr<TupleSetIndex>& getTupleSetIndex() const { return tupleSetIndex_; }
  void clear() { tupleSetIndex_.reset(); }

 private:
  std::vector<std::string> columnNames_;
  std::shared_ptr<TupleSetIndex> tupleSetIndex_;
};

class HashJoinBuildPOp : public PhysicalOp {
 public:
  HashJoinBuildPOp(const std::string& name, const std::vector<std::string>& projectColumnNames, int nodeId,
                   const std::vector<std::string>& predColumnNames);
  std::string getTypeString() const override { return "HashJoinBuildPOp"; }
  void onReceive(const Envelope& msg) override;
  void clear() override { kernel_.clear(); }

 private:
  void send(bool force);
  HashJoinBuildKernel kernel_;
};

class HashJoinProbeKernel {
 public:
  HashJoinProbeKernel(HashJoinPredicate pred, std::set<std::string> neededColumnNames, JoinType joinType);

  Expected<void> joinBuildTupleSetIndex(const std::shared_ptr<TupleSetIndex>& tupleSetIndex);
  Expected<void> joinProbeTupleSet(const std::shared_ptr<TupleSet>& tupleSet);
  // Outer rows (LEFT, RIGHT, FULL) or the semi join's rows (LEFT_SEMI).
  Expected<void> finalize();

  const std::optional<Schema>& getOutputSchema() const { return outputSchema_; }
  const std::optional<std::shared_ptr<TupleSet>>& getBuffer() const { return buffer_; }
  void clearBuffer() { buffer_.reset(); }
  void clear();

 private:
  Expected<void> putBuildTupleSetIndex(const std::shared_ptr<TupleSetIndex>& tupleSetIndex);
  Expected<void> putProbeTupleSet(const std::shared_ptr<TupleSet>& tupleSet);
  Expected<void> bufferOutputSchema(const Schema& build, const Schema& probe);
  Expected<void> joinPairs(TupleSetIndex& build, const std::shared_ptr<TupleSet>& probe);
  Expected<std::vector<Column>> gatherSide(const TupleSet& in, const std::vector<int>& cols, const int64_t* d_rows,
                                           int64_t count, size_t firstField);
  Expected<void> buffer(const std::shared_ptr<TupleSet>& tupleSet);

  HashJoinPredicate pred_;
  std::set<std::string> neededColumnNames_;
  JoinType joinType_;
  std::optional<Schema> outputSchema_;
  std::vector<int> buildOut_, probeOut_;        // needed columns of each side, in output order
  std::shared_ptr<TupleSetIndex> buildIndex_;   // all build rows received
  std::shared_ptr<TupleSet> probeTupleSet_;     // all probe rows received
  std::optional<std::shared_ptr<TupleSet>> buffer_;
};

class HashJoinProbePOp : public PhysicalOp {
 public:
  HashJoinProbePOp(const std::string& name, const std::vector<std::string>& projectColumnNames, int nodeId,
                   const HashJoinPredicate& pred, JoinType joinType);
  std::string getTypeString() const override { return "HashJoinProbePOp"; }
  void onReceive(const Envelope& msg) override;
  void clear() override { kernel_.clear(); }

 private:
  void send(bool force);
  void sendEmpty();
  HashJoinProbeKernel kernel_;
  bool sentResult_ = false;
};

}  // namespace fpdb::amd
