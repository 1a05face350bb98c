// Hash join operator: HashJoinArrowPOp + HashJoinArrowKernel
// (fpdb-executor/include/fpdb/executor/physical/join/hashjoin/HashJoinArrowPOp.h:20-41,
// HashJoinArrowKernel.h:30-99) over the device hash table of the C ABI.
#pragma once

#include <optional>
#include <set>

#include "KeyDictionary.h"
#include "PhysicalOp.h"

namespace fpdb::amd {

// fpdb::plan::prephysical::JoinType (fpdb-plan/include/fpdb/plan/prephysical/JoinType.h:13-20)
enum JoinType { INNER, LEFT, RIGHT, FULL, LEFT_SEMI, RIGHT_SEMI };

// Left columns belong to the build side, right columns to the probe side
// (HashJoinPredicate.cpp:13-68).
struct HashJoinPredicate {
  std::vector<std::string> leftColumnNames;
  std::vector<std::string> rightColumnNames;
};

class HashJoinArrowKernel {
 public:
  HashJoinArrowKernel(HashJoinPredicate pred, std::set<std::string> neededColumnNames, JoinType joinType);
  ~HashJoinArrowKernel();
  HashJoinArrowKernel(const HashJoinArrowKernel&) = delete;
  HashJoinArrowKernel& operator=(const HashJoinArrowKernel&) = delete;

  Expected<void> joinBuildTupleSet(const std::shared_ptr<TupleSet>& tupleSet);
  Expected<void> joinProbeTupleSet(const std::shared_ptr<TupleSet>& tupleSet);
  Expected<void> finalizeInput(bool isBuildSide);

  const std::optional<Schema>& getOutputSchema() const { return outputSchema_; }
  // Output ready to send: at least DefaultBufferSize rows, or the rest once both
  // sides are finalized (BufferedSinkNode, fpdb-tuple/src/arrow/exec/BufferedSinkNode.cpp:32-40).
  const std::optional<std::shared_ptr<TupleSet>>& getOutputBuffer() const { return outputBuffer_; }
  void clearOutputBuffer() { outputBuffer_.reset(); }
  void clear();
  JoinType getJoinType() const { return joinType_; }

 private:
  Expected<void> makeOutputSchema();
  Expected<void> buildTable();
  Expected<void> probe(const std::shared_ptr<TupleSet>& tupleSet);
  Expected<void> finish();
  Expected<void> bufferOutput(const std::shared_ptr<TupleSet>& ts, bool flush);

  HashJoinPredicate pred_;
  std::set<std::string> neededColumnNames_;
  JoinType joinType_;
  std::optional<Schema> buildSchema_, probeSchema_, outputSchema_;
  std::vector<int> probeOut_, buildOut_;     // indices of output columns in each input schema
  std::vector<std::shared_ptr<TupleSet>> buildInput_, pendingProbe_;
  std::shared_ptr<TupleSet> buildTable_;      // concatenated build side
  fpdb_hash_table* table_ = nullptr;
  int payloadColumn_ = -1;                    // build column stored inline in the table
  std::unique_ptr<KeyDictionary> keyDict_;    // keys the join kernels do not take directly
  Column buildKeyIds_;                        // the build keys' ids while the table lives
  Column buildRows_;                          // build row numbers, for utf8 build outputs
  bool buildFinalized_ = false, probeFinalized_ = false, finished_ = false;
  std::shared_ptr<TupleSet> pendingOutput_;
  std::optional<std::shared_ptr<TupleSet>> outputBuffer_;
};

class HashJoinArrowPOp : public PhysicalOp {
 public:
  HashJoinArrowPOp(const std::string& name, const std::vector<std::string>& projectColumnNames, int nodeId,
                   const HashJoinPredicate& pred, JoinType joinType);
  std::string getTypeString() const override { return "HashJoinArrowPOp"; }
  void onReceive(const Envelope& msg) override;
  void clear() override { kernel_.clear(); }
  void addBuildProducer(const std::shared_ptr<PhysicalOp>& buildProducer);
  void addProbeProducer(const std::shared_ptr<PhysicalOp>& probeProducer);
  const HashJoinArrowKernel& getKernel() const { return kernel_; }
  int64_t numRowsBuild() const { return numRowsBuild_; }
  int64_t numRowsProbe() const { return numRowsProbe_; }

 private:
  void onTupleSet(const TupleSetMessage& msg);
  void onComplete(const CompleteMessage& msg);
  void send();
  void sendEmpty();

  HashJoinArrowKernel kernel_;
  std::set<std::string> buildProducers_, probeProducers_;
  int numCompletedBuildProducers_ = 0, numCompletedProbeProducers_ = 0;
  bool sentResult_ = false;
  int64_t numRowsBuild_ = 0, numRowsProbe_ = 0;
};

}  // namespace fpdb::amd
