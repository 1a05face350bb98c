// Hash group-by operator: GroupPOp + GroupArrowKernel
// (fpdb-executor/include/fpdb/executor/physical/group/GroupPOp.h:36-40,
// GroupArrowKernel.h) and the aggregate functions it evaluates
// (fpdb-executor/src/physical/aggregate/function/{Sum,Count,MinMax,Avg}.cpp).
#pragma once

#include <optional>

#include "KeyDictionary.h"
#include "PhysicalOp.h"

namespace fpdb::amd {

enum AggregateFunctionType { SUM, COUNT, MIN_MAX, AVG, AVG_REDUCE };

// Intermediate column prefixes of a two-phase AVG
// (fpdb-plan/include/fpdb/plan/prephysical/AggregatePrePFunction.h:25-26).
constexpr const char* kAvgIntermediateSumPrefix = "AVG_INTERMEDIATE_SUM_";
constexpr const char* kAvgIntermediateCountPrefix = "AVG_INTERMEDIATE_COUNT_";

// An aggregate over one input column (the reference evaluates a gandiva
// expression; here the expression is a column reference, empty for count(*)).
class AggregateFunction {
 public:
  AggregateFunction(AggregateFunctionType type, std::string outputColumnName, std::string inputColumnName)
      : type_(type), outputColumnName_(std::move(outputColumnName)), inputColumnName_(std::move(inputColumnName)) {}
  virtual ~AggregateFunction() = default;
  AggregateFunctionType getType() const { return type_; }
  const std::string& getOutputColumnName() const { return outputColumnName_; }
  const std::string& getInputColumnName() const { return inputColumnName_; }
  // AGG_INPUT_<output> naming of AggregateFunction.cpp:199-201
  std::string getAggregateInputColumnName() const { return "AGG_INPUT_" + outputColumnName_; }
  virtual fpdb_agg_kind kind() const = 0;

 protected:
  AggregateFunctionType type_;
  std::string outputColumnName_;
  std::string inputColumnName_;
};

struct Sum : AggregateFunction {
  Sum(std::string out, std::string in) : AggregateFunction(SUM, std::move(out), std::move(in)) {}
  fpdb_agg_kind kind() const override { return FPDB_AGG_SUM; }
};

// count(*) when inputColumnName is empty (hash_count ALL), else ONLY_VALID
// (Count.cpp:91-104).
struct Count : AggregateFunction {
  Count(std::string out, std::string in = "") : AggregateFunction(COUNT, std::move(out), std::move(in)) {}
  fpdb_agg_kind kind() const override { return inputColumnName_.empty() ? FPDB_AGG_COUNT_ALL : FPDB_AGG_COUNT_VALID; }
};

struct MinMax : AggregateFunction {
  MinMax(bool isMin, std::string out, std::string in)
      : AggregateFunction(MIN_MAX, std::move(out), std::move(in)), isMin_(isMin) {}
  fpdb_agg_kind kind() const override { return isMin_ ? FPDB_AGG_MIN : FPDB_AGG_MAX; }
  bool isMin_;
};

struct Avg : AggregateFunction {
  Avg(std::string out, std::string in) : AggregateFunction(AVG, std::move(out), std::move(in)) {}
  fpdb_agg_kind kind() const override { return FPDB_AGG_MEAN; }
};

// The reducing half of a two-phase AVG (AvgReduce.cpp:12-115): sums the
// partial operators' intermediate sum and count columns per group and divides
// them (AvgBase::finalize, AvgBase.cpp:71-98).  Its inputs are the columns
// AVG_INTERMEDIATE_SUM_<out> and AVG_INTERMEDIATE_COUNT_<out>.
struct AvgReduce : AggregateFunction {
  explicit AvgReduce(std::string out) : AggregateFunction(AVG_REDUCE, std::move(out), "") {}
  std::string getIntermediateSumColumnName() const { return kAvgIntermediateSumPrefix + outputColumnName_; }
  std::string getIntermediateCountColumnName() const { return kAvgIntermediateCountPrefix + outputColumnName_; }
  fpdb_agg_kind kind() const override { return FPDB_AGG_SUM; }
};

// The partial half of a two-phase AVG over `in` (PrePToPTransformerUtil.cpp:79-87):
// a Sum into AVG_INTERMEDIATE_SUM_<out> and a Count into
// AVG_INTERMEDIATE_COUNT_<out>, which the reducing AvgReduce(out) reads.
inline std::vector<std::shared_ptr<AggregateFunction>> makeAvgPartial(const std::string& out, const std::string& in) {
  return {std::make_shared<Sum>(kAvgIntermediateSumPrefix + out, in),
          std::make_shared<Count>(kAvgIntermediateCountPrefix + out, in)};
}

class GroupArrowKernel {
 public:
  GroupArrowKernel(std::vector<std::string> groupColumnNames,
                   std::vector<std::shared_ptr<AggregateFunction>> aggregateFunctions);
  ~GroupArrowKernel();
  GroupArrowKernel(const GroupArrowKernel&) = delete;
  GroupArrowKernel& operator=(const GroupArrowKernel&) = delete;

  Expected<void> group(const std::shared_ptr<TupleSet>& tupleSet);
  // [group columns..., aggregates in function order] (GroupArrowKernel::finalizeAvg,
  // GroupArrowKernel.cpp:156-200), groups in first-appearance order.
  Expected<std::shared_ptr<TupleSet>> finalise();
  void clear();

 private:
  Expected<void> makeOutputSchema(const Schema& in);
  Expected<void> consume(size_t input);
  Expected<void> switchToDictionary();
  Expected<void> recreate(int64_t maxGroups);
  Expected<bool> widenKeyRanges(const TupleSet& ts);

  std::vector<std::string> groupColumnNames_;
  std::vector<std::shared_ptr<AggregateFunction>> aggregateFunctions_;
  std::optional<Schema> outputSchema_;
  std::vector<int32_t> keyTypes_;
  // one device aggregate per function, two for an AvgReduce (sum of the
  // intermediate sums, then sum of the intermediate counts)
  std::vector<fpdb_agg_spec> specs_;
  std::vector<std::string> specInputs_;        // input column of each device aggregate ("" for count(*))
  std::vector<int> firstSpecOf_;               // per function: its first device aggregate
  std::vector<std::shared_ptr<TupleSet>> inputs_;   // kept to rebuild a table that overflowed
  fpdb_group_table* table_ = nullptr;
  int64_t maxGroups_ = 1 << 16;
  // Keys that do not fit the plain 64-bit key (more than two columns, or two
  // with one wider than 4 bytes) are packed from their value ranges, which
  // grow as batches arrive.
  bool packed_ = false;
  bool haveRange_ = false;
  std::vector<int64_t> keyMin_, keyMax_;     // value range seen so far
  std::vector<int64_t> packMin_;             // packing of the current table
  std::vector<int32_t> packBits_;
  // Keys of other kinds -- utf8, nullable (null is a group, as in Arrow's
  // Grouper), or too wide to pack into 64 bits -- are grouped by their
  // key-dictionary id (one INT32 key); the dictionary turns ids back into keys.
  bool useDict_ = false;
  std::unique_ptr<KeyDictionary> keyDict_;
  std::vector<Column> inputIds_;             // ids of inputs_[i], kept for rebuilds
};

class GroupPOp : public PhysicalOp {
 public:
  GroupPOp(const std::string& name, const std::vector<std::string>& projectColumnNames, int nodeId,
           const std::vector<std::string>& groupColumnNames,
           const std::vector<std::shared_ptr<AggregateFunction>>& aggregateFunctions);
  std::string getTypeString() const override { return "GroupPOp"; }
  void onReceive(const Envelope& msg) override;
  void clear() override { kernel_.clear(); }

 private:
  void onTupleSet(const TupleSetMessage& msg);
  void onComplete(const CompleteMessage& msg);
  GroupArrowKernel kernel_;
};

}  // namespace fpdb::amd
