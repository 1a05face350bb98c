// Plan plumbing around the hot-path operators, so plans can run end to end in
// one process: an in-memory source (in place of the reference's scans), a
// column-vs-constant filter (in place of gandiva FilterPOp) and a collate sink
// (CollatePOp, fpdb-executor/src/physical/collate/CollatePOp.cpp).
#pragma once

#include "PhysicalOp.h"

namespace fpdb::amd {

class SourcePOp : public PhysicalOp {
 public:
  SourcePOp(const std::string& name, std::vector<std::shared_ptr<TupleSet>> batches,
            std::vector<std::string> projectColumnNames = {})
      : PhysicalOp(name, POpType::SOURCE, std::move(projectColumnNames), 0), batches_(std::move(batches)) {}
  std::string getTypeString() const override { return "SourcePOp"; }
  void onReceive(const Envelope& msg) override;
  void clear() override { batches_.clear(); }

 private:
  std::vector<std::shared_ptr<TupleSet>> batches_;
};

// Keeps the rows where `column op constant` holds (nulls dropped).
class FilterPOp : public PhysicalOp {
 public:
  FilterPOp(const std::string& name, const std::vector<std::string>& projectColumnNames, int nodeId,
            std::string column, fpdb_cmp_op op, int64_t intConstant, double floatConstant = 0.0)
      : PhysicalOp(name, POpType::FILTER, projectColumnNames, nodeId), column_(std::move(column)), op_(op),
        intConstant_(intConstant), floatConstant_(floatConstant) {}
  // `column op stringConstant` on a utf8 column (bytewise order)
  FilterPOp(const std::string& name, const std::vector<std::string>& projectColumnNames, int nodeId,
            std::string column, fpdb_cmp_op op, std::string stringConstant)
      : PhysicalOp(name, POpType::FILTER, projectColumnNames, nodeId), column_(std::move(column)), op_(op),
        intConstant_(0), floatConstant_(0.0), hasString_(true), stringConstant_(std::move(stringConstant)) {}
  std::string getTypeString() const override { return "FilterPOp"; }
  void onReceive(const Envelope& msg) override;
  void clear() override {}
  static Expected<std::shared_ptr<TupleSet>> filter(const std::shared_ptr<TupleSet>& ts, const std::string& column,
                                                    fpdb_cmp_op op, int64_t intConstant, double floatConstant,
                                                    const std::string* stringConstant = nullptr);

 private:
  std::string column_;
  fpdb_cmp_op op_;
  int64_t intConstant_;
  double floatConstant_;
  bool hasString_ = false;
  std::string stringConstant_;
};

class CollatePOp : public PhysicalOp {
 public:
  explicit CollatePOp(const std::string& name) : PhysicalOp(name, POpType::COLLATE, {}, 0) {}
  std::string getTypeString() const override { return "CollatePOp"; }
  void onReceive(const Envelope& msg) override;
  void clear() override { received_.clear(); }
  Expected<std::shared_ptr<TupleSet>> tupleSet() const;

 private:
  std::vector<std::shared_ptr<TupleSet>> received_;
};

}  // namespace fpdb::amd
