#include "Group.h"

namespace fpdb::amd {

GroupArrowKernel::GroupArrowKernel(std::vector<std::string> groupColumnNames,
                                   std::vector<std::shared_ptr<AggregateFunction>> aggregateFunctions)
    : groupColumnNames_(std::move(groupColumnNames)), aggregateFunctions_(std::move(aggregateFunctions)) {}

GroupArrowKernel::~GroupArrowKernel() { clear(); }

void GroupArrowKernel::clear() {
  if (table_) fpdb_groupby_destroy(table_);
  table_ = nullptr;
  inputs_.clear();
  inputIds_.clear();
  keyDict_.reset();
  haveRange_ = false;
  keyMin_.clear();
  keyMax_.clear();
  packMin_.clear();
  packBits_.clear();
}

Expected<void> GroupArrowKernel::makeOutputSchema(const Schema& in) {
  if (outputSchema_.has_value()) return {};
  Schema out;
  keyTypes_.clear();
  specs_.clear();
  for (auto& g : groupColumnNames_) {
    int i = in.getFieldIndex(g);
    if (i < 0) return make_unexpected("Group field '" + g + "' not found in input schema");
    out.fields.push_back(in.fields[i]);
    keyTypes_.push_back(in.fields[i].type);
  }
  packed_ = keyTypes_.size() > 2;
  if (keyTypes_.size() == 2)
    for (auto t : keyTypes_) packed_ = packed_ || typeWidth((fpdb_type)t) > 4;
  for (auto& g : groupColumnNames_) {
    const Field& f = in.fields[in.getFieldIndex(g)];
    if (f.type == FPDB_TYPE_UTF8 || f.nullable) useDict_ = true;
  }
  if (useDict_) packed_ = false;
  specInputs_.clear();
  firstSpecOf_.clear();
  auto addSpec = [&](fpdb_agg_kind kind, const std::string& input) -> Expected<fpdb_agg_spec> {
    fpdb_agg_spec s{};
    s.kind = kind;
    if (!input.empty()) {
      int i = in.getFieldIndex(input);
      if (i < 0) return make_unexpected("Aggregate input '" + input + "' not found");
      s.input_type = in.fields[i].type;
      s.input_nullable = in.fields[i].nullable ? 1 : 0;
    } else {
      s.input_type = FPDB_TYPE_INT64;
    }
    specs_.push_back(s);
    specInputs_.push_back(input);
    return s;
  };
  for (auto& f : aggregateFunctions_) {
    firstSpecOf_.push_back((int)specs_.size());
    Field of;
    of.name = f->getOutputColumnName();
    if (f->getType() == AVG_REDUCE) {
      // hash_sum of both intermediate columns (AvgReduce.cpp:84-106); the
      // average is their quotient, float64 (AvgBase.cpp:71-98)
      auto ar = std::static_pointer_cast<AvgReduce>(f);
      auto s1 = addSpec(FPDB_AGG_SUM, ar->getIntermediateSumColumnName());
      if (!s1) return make_unexpected(s1.error());
      auto s2 = addSpec(FPDB_AGG_SUM, ar->getIntermediateCountColumnName());
      if (!s2) return make_unexpected(s2.error());
      of.type = FPDB_TYPE_FLOAT64;
      of.nullable = true;
    } else {
      auto s = addSpec(f->kind(), f->getInputColumnName());
      if (!s) return make_unexpected(s.error());
      of.type = (fpdb_type)fpdb_groupby_output_type(*s);
      of.nullable = s->kind != FPDB_AGG_COUNT_ALL && s->kind != FPDB_AGG_COUNT_VALID;
    }
    out.fields.push_back(of);
  }
  outputSchema_ = out;
  return {};
}

Expected<void> GroupArrowKernel::recreate(int64_t maxGroups) {
  if (table_) fpdb_groupby_destroy(table_);
  table_ = nullptr;
  maxGroups_ = maxGroups;
  if (useDict_) {
    const int32_t idType = FPDB_TYPE_INT32;
    if (fpdb_groupby_create(defaultContext(), &idType, 1, specs_.data(), (int)specs_.size(), maxGroups_, &table_) !=
        FPDB_OK)
      return make_unexpected(fpdb_last_error());
    return {};
  }
  fpdb_status s = packed_ ? fpdb_groupby_create_packed(defaultContext(), keyTypes_.data(), (int)keyTypes_.size(),
                                                       packMin_.data(), packBits_.data(), specs_.data(),
                                                       (int)specs_.size(), maxGroups_, &table_)
                          : fpdb_groupby_create(defaultContext(), keyTypes_.data(), (int)keyTypes_.size(),
                                                specs_.data(), (int)specs_.size(), maxGroups_, &table_);
  if (s != FPDB_OK) return make_unexpected(fpdb_last_error());
  return {};
}

static int bitLength(uint64_t v) {
  int b = 0;
  while (v) {
    ++b;
    v >>= 1;
  }
  return b;
}

// Widens the key value ranges by one batch; true when the current table's
// packing no longer covers them (a new packing is then chosen: each key gets
// the bits of its span, plus one spare bit each while the total stays <= 64,
// so a slowly growing key does not force a rebuild per batch).
Expected<bool> GroupArrowKernel::widenKeyRanges(const TupleSet& ts) {
  size_t nk = groupColumnNames_.size();
  for (size_t j = 0; j < nk; ++j) {
    fpdb_column col = ts.column(ts.schema().getFieldIndex(groupColumnNames_[j])).c();
    int64_t lo = 0, hi = 0;
    if (fpdb_column_min_max(defaultContext(), &col, ts.numRows(), &lo, &hi, nullptr) != FPDB_OK)
      return make_unexpected(fpdb_last_error());
    if (!haveRange_) {
      keyMin_.push_back(lo);
      keyMax_.push_back(hi);
    } else {
      keyMin_[j] = std::min(keyMin_[j], lo);
      keyMax_[j] = std::max(keyMax_[j], hi);
    }
  }
  haveRange_ = true;
  bool fits = table_ != nullptr && packMin_.size() == nk;
  for (size_t j = 0; fits && j < nk; ++j) {
    uint64_t span = (uint64_t)keyMax_[j] - (uint64_t)packMin_[j];
    fits = keyMin_[j] >= packMin_[j] && (packBits_[j] >= 64 || span < (1ULL << packBits_[j]));
  }
  if (fits) return false;
  std::vector<int32_t> bits(nk);
  int total = 0;
  for (size_t j = 0; j < nk; ++j) {
    bits[j] = std::max(1, bitLength((uint64_t)keyMax_[j] - (uint64_t)keyMin_[j]));
    total += bits[j];
  }
  if (total > 64) return make_unexpected("Group keys need more than 64 bits when packed");
  for (size_t j = 0; j < nk && total < 64; ++j, ++total) bits[j] += 1;
  packMin_ = keyMin_;
  packBits_ = bits;
  return true;
}

// Packed keys that outgrow 64 bits: group by key-dictionary ids from here on
// (the batches so far are encoded and consumed again).
Expected<void> GroupArrowKernel::switchToDictionary() {
  useDict_ = true;
  packed_ = false;
  keyDict_.reset();
  inputIds_.clear();
  return {};
}

Expected<void> GroupArrowKernel::consume(size_t input) {
  const auto& ts = inputs_[input];
  std::vector<fpdb_column> keys, vals;
  for (auto& g : groupColumnNames_) keys.push_back(ts->column(ts->schema().getFieldIndex(g)).c());
  if (useDict_) {
    if (!keyDict_) {
      std::vector<fpdb_type> kt;
      for (auto t : keyTypes_) kt.push_back((fpdb_type)t);
      try {
        keyDict_ = std::make_unique<KeyDictionary>(kt);
      } catch (const std::exception& e) {
        return make_unexpected(e.what());
      }
    }
    if (inputIds_.size() <= input) {
      auto ids = keyDict_->encode(keys, ts->numRows(), true, true);
      if (!ids) return make_unexpected(ids.error());
      inputIds_.push_back(*ids);
    }
    // null_is_key ids are never null: the group-by takes them without a bitmap
    fpdb_column ids = inputIds_[input].c();
    ids.validity = nullptr;
    keys = {ids};
  }
  for (auto& input : specInputs_) {
    if (input.empty()) {
      vals.push_back(fpdb_column{});
    } else {
      vals.push_back(ts->column(ts->schema().getFieldIndex(input)).c());
    }
  }
  if (fpdb_groupby_consume(defaultContext(), table_, keys.data(), vals.data(), ts->numRows()) != FPDB_OK)
    return make_unexpected(fpdb_last_error());
  return {};
}

// GroupArrowKernel::group (GroupArrowKernel.cpp:16-39).  Batches are consumed as
// they arrive; if the table turns out too small, it is recreated larger and
// the batches seen so far are consumed again in arrival order.
Expected<void> GroupArrowKernel::group(const std::shared_ptr<TupleSet>& tupleSet) {
  auto r = makeOutputSchema(tupleSet->schema());
  if (!r) return r;
  if (tupleSet->numRows() == 0) return {};
  bool rebuild = !table_;
  if (packed_) {
    auto w = widenKeyRanges(*tupleSet);
    if (w) {
      rebuild = rebuild || *w;
    } else {
      r = switchToDictionary();
      if (!r) return r;
      rebuild = true;
    }
  }
  inputs_.push_back(tupleSet);
  if (rebuild) {
    r = recreate(std::max<int64_t>(maxGroups_, 1));
    if (!r) return r;
    for (size_t i = 0; i < inputs_.size(); ++i) {
      r = consume(i);
      if (!r) return r;
    }
  } else {
    r = consume(inputs_.size() - 1);
    if (!r) return r;
  }
  int64_t groups = 0;
  fpdb_status s = fpdb_groupby_num_groups(defaultContext(), table_, &groups);
  while (s == FPDB_ERR_CAPACITY) {
    r = recreate(maxGroups_ * 8);
    if (!r) return r;
    for (size_t i = 0; i < inputs_.size(); ++i) {
      r = consume(i);
      if (!r) return r;
    }
    s = fpdb_groupby_num_groups(defaultContext(), table_, &groups);
  }
  if (s != FPDB_OK) return make_unexpected(fpdb_last_error());
  return {};
}

Expected<std::shared_ptr<TupleSet>> GroupArrowKernel::finalise() {
  if (!outputSchema_.has_value()) return make_unexpected("Output schema not made yet");
  const Schema& os = *outputSchema_;
  if (!table_) return TupleSet::make(os);   // no input rows
  fpdb_context* ctx = defaultContext();
  int64_t ng = 0;
  if (fpdb_groupby_num_groups(ctx, table_, &ng) != FPDB_OK) return make_unexpected(fpdb_last_error());
  std::vector<Column> cols, aggCols;
  std::vector<fpdb_column> outKeys, outAggs;
  size_t nk = groupColumnNames_.size();
  Column groupIds;
  if (useDict_) {
    groupIds = Column::allocate(ctx, FPDB_TYPE_INT32, ng, false);
    outKeys.push_back(groupIds.c());
  } else {
    for (size_t i = 0; i < nk; ++i) {
      cols.push_back(Column::allocate(ctx, os.fields[i].type, ng, os.fields[i].nullable));
      outKeys.push_back(cols.back().c());
    }
  }
  for (auto& s : specs_) {
    const bool nullable = s.kind != FPDB_AGG_COUNT_ALL && s.kind != FPDB_AGG_COUNT_VALID;
    aggCols.push_back(Column::allocate(ctx, (fpdb_type)fpdb_groupby_output_type(s), ng, nullable));
    outAggs.push_back(aggCols.back().c());
  }
  int64_t written = 0;
  if (fpdb_groupby_finalize(ctx, table_, outKeys.data(), outAggs.data(), &written) != FPDB_OK)
    return make_unexpected(fpdb_last_error());
  if (useDict_) {
    // each group's key values from its id
    for (size_t i = 0; i < nk; ++i) {
      auto k = keyDict_->decode((int)i, groupIds, ng);
      if (!k) return make_unexpected(k.error());
      cols.push_back(*k);
    }
  }
  // [group columns..., aggregates in function order]; an AvgReduce divides its
  // two sums (GroupArrowKernel::finalizeAvg, GroupArrowKernel.cpp:156-200)
  for (size_t f = 0; f < aggregateFunctions_.size(); ++f) {
    const int s = firstSpecOf_[f];
    if (aggregateFunctions_[f]->getType() != AVG_REDUCE) {
      cols.push_back(aggCols[s]);
      continue;
    }
    Column avg = Column::allocate(ctx, FPDB_TYPE_FLOAT64, ng, true);
    fpdb_column num = aggCols[s].c(), den = aggCols[s + 1].c(), o = avg.c();
    if (fpdb_divide_to_float64(ctx, &num, &den, ng, &o) != FPDB_OK) return make_unexpected(fpdb_last_error());
    cols.push_back(avg);
  }
  return TupleSet::make(os, cols);
}

GroupPOp::GroupPOp(const std::string& name, const std::vector<std::string>& projectColumnNames, int nodeId,
                   const std::vector<std::string>& groupColumnNames,
                   const std::vector<std::shared_ptr<AggregateFunction>>& aggregateFunctions)
    : PhysicalOp(name, POpType::GROUP, projectColumnNames, nodeId), kernel_(groupColumnNames, aggregateFunctions) {}

void GroupPOp::onReceive(const Envelope& msg) {
  switch (msg.message().type()) {
    case MessageType::START: break;
    case MessageType::TUPLESET: onTupleSet(static_cast<const TupleSetMessage&>(msg.message())); break;
    case MessageType::COMPLETE: onComplete(static_cast<const CompleteMessage&>(msg.message())); break;
    default: ctx()->notifyError("Unrecognized message type " + msg.message().getTypeString());
  }
}

void GroupPOp::onTupleSet(const TupleSetMessage& message) {
  auto r = kernel_.group(message.tuples());
  if (!r) ctx()->notifyError(r.error());
}

// GroupPOp.cpp:67-85: finalise, project, send, complete
void GroupPOp::onComplete(const CompleteMessage&) {
  if (!ctx()->isComplete() && ctx()->operatorMap().allComplete(POpRelationshipType::Producer)) {
    auto grouped = kernel_.finalise();
    if (!grouped) {
      ctx()->notifyError(grouped.error());
      return;
    }
    auto projected = (*grouped)->projectExist(getProjectColumnNames());
    if (!projected) {
      ctx()->notifyError(projected.error());
      return;
    }
    ctx()->tell(std::make_shared<TupleSetMessage>(*projected, name_));
    ctx()->notifyComplete();
  }
}

}  // namespace fpdb::amd
