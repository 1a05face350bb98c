// Predicate-transfer operators: BloomFilterCreatePOp / BloomFilterUsePOp
// (fpdb-executor/include/fpdb/executor/physical/bloomfilter/BloomFilterCreatePOp.h:19-40,
// BloomFilterUsePOp.h:17-20) and their kernels.
#pragma once

#include <optional>
#include <set>

#include "BloomFilter.h"
#include "PhysicalOp.h"

namespace fpdb::amd {

// BloomFilterCreateAbstractKernel (BloomFilterCreateAbstractKernel.cpp:22-38)
// with its Arrow (blocked) and legacy builds.
class BloomFilterCreateKernel {
 public:
  BloomFilterCreateKernel(std::vector<std::string> columnNames, double desiredFalsePositiveRate, bool arrowImpl);
  // Buffers input; the reference never mutates it (BloomFilterCreateAbstractKernel.cpp:29-31).
  Expected<void> bufferTupleSet(const std::shared_ptr<TupleSet>& tupleSet);
  Expected<void> buildBloomFilter();
  std::optional<std::shared_ptr<BloomFilterBase>> getBloomFilter() const { return bloomFilter_; }
  void clear();

 private:
  std::vector<std::string> columnNames_;
  double fpr_;
  bool arrowImpl_;
  std::vector<std::shared_ptr<TupleSet>> received_;
  std::optional<std::shared_ptr<BloomFilterBase>> bloomFilter_;
};

// BloomFilterUseKernel::filter (BloomFilterUseKernel.cpp:12-32): the rows of
// the TupleSet whose key may be in the filter, all columns kept, in order.
struct BloomFilterUseKernel {
  static Expected<std::shared_ptr<TupleSet>> filter(const std::shared_ptr<TupleSet>& tupleSet,
                                                    const std::shared_ptr<BloomFilterBase>& bloomFilter,
                                                    const std::vector<std::string>& columnNames);
};

class BloomFilterCreatePOp : public PhysicalOp {
 public:
  BloomFilterCreatePOp(const std::string& name, const std::vector<std::string>& projectColumnNames, int nodeId,
                       const std::vector<std::string>& bloomFilterColumnNames,
                       double desiredFalsePositiveRate = 0.01);
  std::string getTypeString() const override { return "BloomFilterCreatePOp"; }
  void onReceive(const Envelope& msg) override;
  void clear() override { kernel_.clear(); }
  void produce(const std::shared_ptr<PhysicalOp>& op) override;
  void addBloomFilterUsePOp(const std::shared_ptr<PhysicalOp>& bloomFilterUsePOp);
  const std::set<std::string>& getBloomFilterUsePOps() const { return bloomFilterUsePOps_; }
  const std::optional<std::shared_ptr<BloomFilterBase>> getBloomFilter() const { return kernel_.getBloomFilter(); }

 private:
  void onTupleSet(const TupleSetMessage& msg);
  void onComplete(const CompleteMessage& msg);

  BloomFilterCreateKernel kernel_;
  std::set<std::string> bloomFilterUsePOps_;
  std::set<std::string> passTupleSetConsumers_;
};

class BloomFilterUsePOp : public PhysicalOp {
 public:
  BloomFilterUsePOp(const std::string& name, const std::vector<std::string>& projectColumnNames, int nodeId,
                    const std::vector<std::string>& bloomFilterColumnNames);
  std::string getTypeString() const override { return "BloomFilterUsePOp"; }
  void onReceive(const Envelope& msg) override;
  void clear() override;
  int64_t numRowsInput() const { return numRowsInput_; }
  int64_t numRowsOutput() const { return numRowsOutput_; }

 private:
  void onTupleSet(const TupleSetMessage& msg);
  void onBloomFilter(const BloomFilterMessage& msg);
  void onComplete(const CompleteMessage& msg);
  Expected<void> filterAndSend();

  std::vector<std::string> bloomFilterColumnNames_;
  std::optional<std::shared_ptr<TupleSet>> receivedTupleSet_;
  std::optional<std::shared_ptr<BloomFilterBase>> bloomFilter_;
  int64_t numRowsInput_ = 0, numRowsOutput_ = 0;
};

}  // namespace fpdb::amd
