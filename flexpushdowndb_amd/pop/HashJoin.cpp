#include "HashJoin.h"

#include "Globals.h"

namespace fpdb::amd {

HashJoinArrowKernel::HashJoinArrowKernel(HashJoinPredicate pred, std::set<std::string> neededColumnNames,
                                         JoinType joinType)
    : pred_(std::move(pred)), neededColumnNames_(std::move(neededColumnNames)), joinType_(joinType) {}

HashJoinArrowKernel::~HashJoinArrowKernel() { clear(); }

void HashJoinArrowKernel::clear() {
  if (table_) fpdb_hj_table_destroy(table_);
  table_ = nullptr;
  keyDict_.reset();
  buildKeyIds_ = Column();
  buildRows_ = Column();
  buildInput_.clear();
  pendingProbe_.clear();
  buildTable_.reset();
  pendingOutput_.reset();
  outputBuffer_.reset();
}

// Probe-side fields first, then build-side fields (HashJoinArrowKernel.cpp:150-162);
// a semi join keeps one side only.
Expected<void> HashJoinArrowKernel::makeOutputSchema() {
  if (outputSchema_.has_value() || !buildSchema_.has_value() || !probeSchema_.has_value()) return {};
  Schema out;
  bool probeSide = joinType_ != LEFT_SEMI;
  bool buildSide = joinType_ != RIGHT_SEMI;
  bool probeNullable = joinType_ == LEFT || joinType_ == FULL;
  bool buildNullable = joinType_ == RIGHT || joinType_ == FULL;
  probeOut_.clear();
  buildOut_.clear();
  if (probeSide) {
    for (size_t i = 0; i < probeSchema_->fields.size(); ++i) {
      if (!neededColumnNames_.count(probeSchema_->fields[i].name)) continue;
      Field f = probeSchema_->fields[i];
      f.nullable = f.nullable || probeNullable;
      out.fields.push_back(f);
      probeOut_.push_back((int)i);
    }
  }
  if (buildSide) {
    for (size_t i = 0; i < buildSchema_->fields.size(); ++i) {
      if (!neededColumnNames_.count(buildSchema_->fields[i].name)) continue;
      Field f = buildSchema_->fields[i];
      f.nullable = f.nullable || buildNullable;
      out.fields.push_back(f);
      buildOut_.push_back((int)i);
    }
  }
  for (auto& n : pred_.leftColumnNames)
    if (buildSchema_->getFieldIndex(n) < 0) return make_unexpected("Build join column '" + n + "' not found");
  for (auto& n : pred_.rightColumnNames)
    if (probeSchema_->getFieldIndex(n) < 0) return make_unexpected("Probe join column '" + n + "' not found");
  outputSchema_ = out;
  return {};
}

Expected<void> HashJoinArrowKernel::joinBuildTupleSet(const std::shared_ptr<TupleSet>& tupleSet) {
  if (!buildSchema_.has_value()) buildSchema_ = tupleSet->schema();
  auto r = makeOutputSchema();
  if (!r) return r;
  if (tupleSet->numRows() == 0) return {};
  if (buildFinalized_) return make_unexpected("build input after the build side was finalized");
  buildInput_.push_back(tupleSet);
  return {};
}

Expected<void> HashJoinArrowKernel::joinProbeTupleSet(const std::shared_ptr<TupleSet>& tupleSet) {
  if (!probeSchema_.has_value()) probeSchema_ = tupleSet->schema();
  auto r = makeOutputSchema();
  if (!r) return r;
  if (tupleSet->numRows() == 0) return {};
  if (table_) return probe(tupleSet);
  pendingProbe_.push_back(tupleSet);
  return {};
}

Expected<void> HashJoinArrowKernel::buildTable() {
  fpdb_context* ctx = defaultContext();
  if (buildInput_.empty()) {
    buildTable_ = TupleSet::make(*buildSchema_);
  } else {
    auto all = TupleSet::concatenate(buildInput_);
    if (!all) return make_unexpected(all.error());
    buildTable_ = *all;
    buildInput_.clear();
  }
  std::vector<fpdb_column> keys;
  std::vector<fpdb_type> buildTypes, probeTypes;
  for (auto& n : pred_.leftColumnNames) {
    keys.push_back(buildTable_->column(buildSchema_->getFieldIndex(n)).c());
    buildTypes.push_back(buildSchema_->fields[buildSchema_->getFieldIndex(n)].type);
  }
  for (auto& n : pred_.rightColumnNames) probeTypes.push_back(probeSchema_->fields[probeSchema_->getFieldIndex(n)].type);
  if (!joinKeysNative(buildTypes) || !joinKeysNative(probeTypes)) {
    // utf8, float or wide composite keys: one int32 id per distinct key tuple
    try {
      keyDict_ = std::make_unique<KeyDictionary>(buildTypes);
    } catch (const std::exception& e) {
      return make_unexpected(e.what());
    }
    auto ids = keyDict_->encode(keys, buildTable_->numRows(), true, false);
    if (!ids) return make_unexpected(ids.error());
    buildKeyIds_ = *ids;
    keys = {buildKeyIds_.c()};
  }
  // A single 4-byte non-null build output column travels inline in the table slots.
  payloadColumn_ = -1;
  if (buildOut_.size() == 1) {
    const Column& c = buildTable_->column(buildOut_[0]);
    if (!c.isUtf8() && typeWidth(c.type) == 4 && !c.validity) payloadColumn_ = buildOut_[0];
  }
  fpdb_column payload;
  if (payloadColumn_ >= 0) payload = buildTable_->column(payloadColumn_).c();
  bool track = joinType_ == LEFT || joinType_ == FULL || joinType_ == LEFT_SEMI;
  fpdb_status s = fpdb_hj_build(ctx, keys.data(), (int)keys.size(), buildTable_->numRows(),
                                payloadColumn_ >= 0 ? &payload : nullptr, track ? 1 : 0, &table_);
  if (s != FPDB_OK) return make_unexpected(fpdb_last_error());
  for (auto& p : pendingProbe_) {
    auto r = probe(p);
    if (!r) return r;
  }
  pendingProbe_.clear();
  return {};
}

Expected<void> HashJoinArrowKernel::probe(const std::shared_ptr<TupleSet>& ts) {
  fpdb_context* ctx = defaultContext();
  int64_t n = ts->numRows();
  std::vector<fpdb_column> keys, pcols, bcols;
  for (auto& name : pred_.rightColumnNames) keys.push_back(ts->column(ts->schema().getFieldIndex(name)).c());
  Column probeIds;
  if (keyDict_) {
    auto ids = keyDict_->encode(keys, n, false, false);
    if (!ids) return make_unexpected(ids.error());
    probeIds = *ids;
    keys = {probeIds.c()};
  }
  // utf8 output columns go through the probe as row numbers (probe rows, or
  // build rows) and are taken afterwards
  Column probeRows;
  try {
    for (int i : probeOut_) {
      if (ts->column(i).isUtf8()) {
        if (!probeRows.values) probeRows = rowNumbers(n);
        pcols.push_back(probeRows.c());
      } else {
        pcols.push_back(ts->column(i).c());
      }
    }
    for (int i : buildOut_) {
      if (i == payloadColumn_) {
        fpdb_column pc{};
        pc.type = buildTable_->column(i).type;
        bcols.push_back(pc);     // values == NULL: the inline payload
      } else if (buildTable_->column(i).isUtf8()) {
        if (!buildRows_.values) buildRows_ = rowNumbers(buildTable_->numRows());
        bcols.push_back(buildRows_.c());
      } else {
        bcols.push_back(buildTable_->column(i).c());
      }
    }
  } catch (const std::exception& e) {
    return make_unexpected(e.what());
  }
  int jt = (int)joinType_;
  int unique = 0;
  if (fpdb_hj_table_info(table_, nullptr, nullptr, nullptr, &unique) != FPDB_OK)
    return make_unexpected(fpdb_last_error());
  int64_t cap = n;
  if (!unique && joinType_ != RIGHT_SEMI && joinType_ != LEFT_SEMI) {
    if (fpdb_hj_probe_count(ctx, table_, jt, keys.data(), (int)keys.size(), n, &cap) != FPDB_OK)
      return make_unexpected(fpdb_last_error());
  }
  if (joinType_ == LEFT_SEMI) cap = 0;
  const Schema& os = *outputSchema_;
  std::vector<Column> outCols;
  std::vector<fpdb_column> out;
  for (size_t i = 0; i < os.fields.size(); ++i) {
    const fpdb_type t = os.fields[i].type == FPDB_TYPE_UTF8 ? FPDB_TYPE_INT64 : os.fields[i].type;
    outCols.push_back(Column::allocate(ctx, t, cap, os.fields[i].nullable));
    out.push_back(outCols.back().c());
  }
  int64_t produced = 0;
  fpdb_status s = fpdb_hj_probe(ctx, table_, jt, keys.data(), (int)keys.size(), n, pcols.data(), (int)pcols.size(),
                                bcols.data(), (int)bcols.size(), out.data(), cap, &produced);
  if (s != FPDB_OK) return make_unexpected(fpdb_last_error());
  if (joinType_ == LEFT_SEMI) return {};
  for (auto& c : outCols) c.length = produced;
  // resolve the utf8 columns from the row numbers the probe wrote
  const size_t nProbe = probeOut_.size();
  for (size_t i = 0; i < os.fields.size(); ++i) {
    if (os.fields[i].type != FPDB_TYPE_UTF8) continue;
    const Column& src = i < nProbe ? ts->column(probeOut_[i]) : buildTable_->column(buildOut_[i - nProbe]);
    auto t = takeUtf8(src, outCols[i].c(), produced);
    if (!t) return make_unexpected(t.error());
    outCols[i] = *t;
  }
  return bufferOutput(TupleSet::make(os, outCols), false);
}

// After both sides: the build rows of LEFT / FULL (unmatched, probe side null)
// and LEFT_SEMI (matched) joins.
Expected<void> HashJoinArrowKernel::finish() {
  if (finished_) return {};
  finished_ = true;
  if (joinType_ == LEFT || joinType_ == FULL || joinType_ == LEFT_SEMI) {
    fpdb_context* ctx = defaultContext();
    int64_t nb = buildTable_->numRows();
    DeviceBuffer rows(ctx, (uint64_t)std::max<int64_t>(nb, 1) * 8);
    int64_t count = 0;
    fpdb_status s = fpdb_hj_build_side_rows(ctx, table_, joinType_ == LEFT_SEMI ? 1 : 0,
                                            static_cast<int64_t*>(rows.data()), std::max<int64_t>(nb, 1), &count);
    if (s != FPDB_OK) return make_unexpected(fpdb_last_error());
    const Schema& os = *outputSchema_;
    std::vector<Column> outCols;
    size_t nProbe = probeOut_.size();
    for (size_t i = 0; i < nProbe; ++i) outCols.push_back(Column::allNull(ctx, os.fields[i].type, count));
    std::vector<fpdb_column> in, out;
    std::vector<Column> gathered;
    fpdb_column rowCol{};
    rowCol.type = FPDB_TYPE_INT64;
    rowCol.length = count;
    rowCol.values = rows.data();
    for (size_t j = 0; j < buildOut_.size(); ++j) {
      const Field& f = os.fields[nProbe + j];
      const Column& src = buildTable_->column(buildOut_[j]);
      if (src.isUtf8()) {
        auto t = takeUtf8(src, rowCol, count);
        if (!t) return make_unexpected(t.error());
        gathered.push_back(*t);
        continue;
      }
      gathered.push_back(Column::allocate(ctx, f.type, count, true));
      in.push_back(src.c());
      out.push_back(gathered.back().c());
    }
    if (!in.empty()) {
      s = fpdb_gather(ctx, in.data(), (int)in.size(), static_cast<int64_t*>(rows.data()), count, out.data());
      if (s != FPDB_OK) return make_unexpected(fpdb_last_error());
    }
    for (auto& g : gathered) {
      g.length = count;
      outCols.push_back(g);
    }
    checkStatus(fpdb_context_synchronize(ctx), "synchronize");
    auto r = bufferOutput(TupleSet::make(os, outCols), false);
    if (!r) return r;
  }
  if (pendingOutput_ && pendingOutput_->numRows() > 0) return bufferOutput(nullptr, true);
  return {};
}

Expected<void> HashJoinArrowKernel::bufferOutput(const std::shared_ptr<TupleSet>& ts, bool flush) {
  if (ts && ts->numRows() > 0) {
    if (!pendingOutput_) {
      pendingOutput_ = std::make_shared<TupleSet>(ts->schema(), ts->columns());
    } else {
      auto r = pendingOutput_->append(ts);
      if (!r) return r;
    }
  }
  if (pendingOutput_ && (flush || pendingOutput_->numRows() >= DefaultBufferSize)) {
    if (!outputBuffer_.has_value()) {
      outputBuffer_ = pendingOutput_;
    } else {
      auto r = (*outputBuffer_)->append(pendingOutput_);
      if (!r) return r;
    }
    pendingOutput_.reset();
  }
  return {};
}

// HashJoinArrowKernel::finalizeInput (HashJoinArrowKernel.cpp:95-118): a side
// is finalized once all its producers completed; an empty side still gets a
// (zero-row) table so outer joins emit the other side.
Expected<void> HashJoinArrowKernel::finalizeInput(bool isBuildSide) {
  if (isBuildSide) {
    if (buildFinalized_) return {};
    buildFinalized_ = true;
    if (!buildSchema_.has_value()) return make_unexpected("build side finalized without a schema");
    if (!outputSchema_.has_value() && probeSchema_.has_value()) {
      auto r = makeOutputSchema();
      if (!r) return r;
    }
    if (!probeSchema_.has_value()) {
      // probe schema unknown yet: build once it is known
      return {};
    }
    auto r = buildTable();
    if (!r) return r;
  } else {
    if (probeFinalized_) return {};
    probeFinalized_ = true;
    if (buildFinalized_ && !table_ && buildSchema_.has_value() && probeSchema_.has_value()) {
      auto r = buildTable();
      if (!r) return r;
    }
  }
  if (buildFinalized_ && probeFinalized_ && table_) return finish();
  return {};
}

// ---- POp -------------------------------------------------------------------------

HashJoinArrowPOp::HashJoinArrowPOp(const std::string& name, const std::vector<std::string>& projectColumnNames,
                                   int nodeId, const HashJoinPredicate& pred, JoinType joinType)
    : PhysicalOp(name, POpType::HASH_JOIN_ARROW, projectColumnNames, nodeId),
      kernel_(pred, std::set<std::string>(projectColumnNames.begin(), projectColumnNames.end()), joinType) {}

void HashJoinArrowPOp::addBuildProducer(const std::shared_ptr<PhysicalOp>& buildProducer) {
  buildProducers_.insert(buildProducer->name());
  consume(buildProducer);
}

void HashJoinArrowPOp::addProbeProducer(const std::shared_ptr<PhysicalOp>& probeProducer) {
  probeProducers_.insert(probeProducer->name());
  consume(probeProducer);
}

void HashJoinArrowPOp::onReceive(const Envelope& msg) {
  switch (msg.message().type()) {
    case MessageType::START: break;
    case MessageType::TUPLESET: onTupleSet(static_cast<const TupleSetMessage&>(msg.message())); break;
    case MessageType::COMPLETE: onComplete(static_cast<const CompleteMessage&>(msg.message())); break;
    default: ctx()->notifyError("Unrecognized message type " + msg.message().getTypeString());
  }
}

// HashJoinArrowPOp.cpp:115-140
void HashJoinArrowPOp::onTupleSet(const TupleSetMessage& message) {
  auto tupleSet = message.tuples();
  const auto& sender = message.sender();
  Expected<void> result;
  if (buildProducers_.count(sender)) {
    numRowsBuild_ += tupleSet->numRows();
    result = kernel_.joinBuildTupleSet(tupleSet);
  } else if (probeProducers_.count(sender)) {
    numRowsProbe_ += tupleSet->numRows();
    result = kernel_.joinProbeTupleSet(tupleSet);
  } else {
    ctx()->notifyError("Unknown sender '" + sender + "', neither build nor probe producer");
    return;
  }
  if (!result) {
    ctx()->notifyError(result.error());
    return;
  }
  send();
}

// HashJoinArrowPOp.cpp:85-113
void HashJoinArrowPOp::onComplete(const CompleteMessage& message) {
  const auto& sender = message.sender();
  Expected<void> r;
  if (buildProducers_.count(sender)) {
    if (++numCompletedBuildProducers_ >= (int)buildProducers_.size()) r = kernel_.finalizeInput(true);
  } else if (probeProducers_.count(sender)) {
    if (++numCompletedProbeProducers_ >= (int)probeProducers_.size()) r = kernel_.finalizeInput(false);
  } else {
    ctx()->notifyError("Unknown sender '" + sender + "', neither build nor probe producer");
    return;
  }
  if (!r) {
    ctx()->notifyError(r.error());
    return;
  }
  if (!ctx()->isComplete() && ctx()->operatorMap().allComplete(POpRelationshipType::Producer)) {
    send();
    if (!sentResult_) sendEmpty();
    ctx()->notifyComplete();
  }
}

void HashJoinArrowPOp::send() {
  const auto& buffer = kernel_.getOutputBuffer();
  if (buffer.has_value()) {
    ctx()->tell(std::make_shared<TupleSetMessage>(*buffer, name_));
    sentResult_ = true;
    kernel_.clearOutputBuffer();
  }
}

// HashJoinArrowPOp.cpp:155-163: an empty result still carries the output schema
void HashJoinArrowPOp::sendEmpty() {
  const auto& schema = kernel_.getOutputSchema();
  if (!schema.has_value()) {
    ctx()->notifyError("Output schema not set yet on completing");
    return;
  }
  ctx()->tell(std::make_shared<TupleSetMessage>(TupleSet::make(*schema), name_));
  sentResult_ = true;
}

}  // namespace fpdb::amd
