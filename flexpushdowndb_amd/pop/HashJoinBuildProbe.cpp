#include "HashJoinBuildProbe.h"

#include <stdexcept>

#include "Globals.h"

namespace fpdb::amd {

// ---- build side ------------------------------------------------------------------

HashJoinBuildKernel::HashJoinBuildKernel(std::vector<std::string> columnNames) : columnNames_(std::move(columnNames)) {}

// HashJoinBuildKernel::put (HashJoinBuildKernel.cpp:24-37)
Expected<void> HashJoinBuildKernel::put(const std::shared_ptr<TupleSet>& tupleSet) {
  if (!tupleSetIndex_) {
    auto idx = TupleSetIndex::make(columnNames_, tupleSet);
    if (!idx) return make_unexpected(idx.error());
    tupleSetIndex_ = *idx;
    return {};
  }
  return tupleSetIndex_->put(tupleSet);
}

HashJoinBuildPOp::HashJoinBuildPOp(const std::string& name, const std::vector<std::string>& projectColumnNames,
                                   int nodeId, const std::vector<std::string>& predColumnNames)
    : PhysicalOp(name, POpType::HASH_JOIN_BUILD, projectColumnNames, nodeId), kernel_(predColumnNames) {
  if (predColumnNames.empty()) throw std::runtime_error("HashJoinBuildPOp '" + name + "': no key columns");
}

// HashJoinBuildPOp.cpp:25-60
void HashJoinBuildPOp::onReceive(const Envelope& msg) {
  switch (msg.message().type()) {
    case MessageType::START: break;
    case MessageType::TUPLESET: {
      auto r = kernel_.put(static_cast<const TupleSetMessage&>(msg.message()).tuples());
      if (!r) {
        ctx()->notifyError(r.error() + ", " + name());
        return;
      }
      send(false);
      break;
    }
    case MessageType::COMPLETE:
      if (!ctx()->isComplete() && ctx()->operatorMap().allComplete(POpRelationshipType::Producer)) {
        send(true);
        ctx()->notifyComplete();
      }
      break;
    default: ctx()->notifyError("Unrecognized message type " + msg.message().getTypeString() + ", " + name());
  }
}

// HashJoinBuildPOp.cpp:66-74: ship the index once it holds DefaultBufferSize
// rows, and whatever is left at completion.
void HashJoinBuildPOp::send(bool force) {
  const auto& idx = kernel_.getTupleSetIndex();
  if (idx && (force || idx->size() >= DefaultBufferSize)) {
    ctx()->tell(std::make_shared<TupleSetIndexMessage>(idx, name()));
    kernel_.clear();
  }
}

// ---- probe side ------------------------------------------------------------------

HashJoinProbeKernel::HashJoinProbeKernel(HashJoinPredicate pred, std::set<std::string> neededColumnNames,
                                         JoinType joinType)
    : pred_(std::move(pred)), neededColumnNames_(std::move(neededColumnNames)), joinType_(joinType) {}

void HashJoinProbeKernel::clear() {
  buildIndex_.reset();
  probeTupleSet_.reset();
  buffer_.reset();
}

static Expected<void> validateColumnNames(const Schema& schema, const std::vector<std::string>& names) {
  for (auto& n : names)
    if (schema.getFieldIndex(n) < 0) return make_unexpected("Column '" + n + "' does not exist.");
  return {};
}

// HashJoinProbeAbstractKernel::putBuildTupleSetIndex (HashJoinProbeAbstractKernel.cpp:13-29).
// The first index is copied rather than kept, so merging later ones never
// changes an index another consumer of the same message holds.
Expected<void> HashJoinProbeKernel::putBuildTupleSetIndex(const std::shared_ptr<TupleSetIndex>& tupleSetIndex) {
  auto v = validateColumnNames(tupleSetIndex->getTupleSet()->schema(), pred_.leftColumnNames);
  if (!v) return make_unexpected("Cannot put build tuple set index into probe kernel. " + v.error());
  if (!buildIndex_) {
    auto idx = TupleSetIndex::make(tupleSetIndex->getColumnNames(), tupleSetIndex->getTupleSet());
    if (!idx) return make_unexpected(idx.error());
    buildIndex_ = *idx;
    return {};
  }
  return buildIndex_->merge(tupleSetIndex);
}

// HashJoinProbeAbstractKernel::putProbeTupleSet (HashJoinProbeAbstractKernel.cpp:31-46)
Expected<void> HashJoinProbeKernel::putProbeTupleSet(const std::shared_ptr<TupleSet>& tupleSet) {
  auto v = validateColumnNames(tupleSet->schema(), pred_.rightColumnNames);
  if (!v) return make_unexpected("Cannot put probe tuple set into probe kernel. " + v.error());
  if (!probeTupleSet_) {
    probeTupleSet_ = std::make_shared<TupleSet>(tupleSet->schema(), tupleSet->columns());
    return {};
  }
  return probeTupleSet_->append(tupleSet);
}

// HashJoinProbeAbstractKernel::bufferOutputSchema (HashJoinProbeAbstractKernel.cpp:60-83):
// the build side's needed fields, then the probe side's.  Columns of the side
// an outer join may leave unmatched become nullable.  A semi join returns
// build rows only, so it takes no probe column.
Expected<void> HashJoinProbeKernel::bufferOutputSchema(const Schema& build, const Schema& probe) {
  if (outputSchema_) return {};
  bool buildNullable = joinType_ == RIGHT || joinType_ == FULL;
  bool probeNullable = joinType_ == LEFT || joinType_ == FULL;
  Schema out;
  buildOut_.clear();
  probeOut_.clear();
  for (size_t i = 0; i < build.fields.size(); ++i) {
    if (!neededColumnNames_.count(build.fields[i].name)) continue;
    Field f = build.fields[i];
    f.nullable = f.nullable || buildNullable;
    out.fields.push_back(f);
    buildOut_.push_back((int)i);
  }
  for (size_t i = 0; i < probe.fields.size(); ++i) {
    if (!neededColumnNames_.count(probe.fields[i].name)) continue;
    if (joinType_ == LEFT_SEMI)
      return make_unexpected("Semi join output takes build-side columns only, not '" + probe.fields[i].name + "'");
    Field f = probe.fields[i];
    f.nullable = f.nullable || probeNullable;
    out.fields.push_back(f);
    probeOut_.push_back((int)i);
  }
  outputSchema_ = out;
  return {};
}

// Matched pairs of an index's rows and probe rows, build columns first: the
// work of RecordBatchHashJoiner::join (RecordBatchHashJoiner.cpp:38-131), one
// fpdb_hj_probe over the whole probe tuple set.
// Key types of the named columns.
static std::vector<fpdb_type> keyTypes(const Schema& schema, const std::vector<std::string>& names) {
  std::vector<fpdb_type> t;
  for (auto& n : names) t.push_back(schema.fields[schema.getFieldIndex(n)].type);
  return t;
}

Expected<void> HashJoinProbeKernel::joinPairs(TupleSetIndex& build, const std::shared_ptr<TupleSet>& probe) {
  if (build.size() == 0 || probe->numRows() == 0) return {};
  auto table = build.table(false, keyTypes(probe->schema(), pred_.rightColumnNames));
  if (!table) return make_unexpected(table.error());
  fpdb_context* ctx = defaultContext();
  int64_t n = probe->numRows();
  std::vector<fpdb_column> rawKeys, cols;
  for (auto& name : pred_.rightColumnNames) rawKeys.push_back(probe->column(probe->schema().getFieldIndex(name)).c());
  Column keyIds;
  auto pk = build.probeKeys(rawKeys, n, &keyIds);
  if (!pk) return make_unexpected(pk.error());
  std::vector<fpdb_column> keys = *pk;
  // utf8 output columns travel as row numbers and are taken afterwards
  Column probeRows, buildRows;
  const TupleSet& bts = *build.getTupleSet();
  try {
    for (int i : probeOut_) {
      if (probe->column(i).isUtf8()) {
        if (!probeRows.values) probeRows = rowNumbers(n);
        cols.push_back(probeRows.c());
      } else {
        cols.push_back(probe->column(i).c());
      }
    }
    for (int i : buildOut_) {
      if (bts.column(i).isUtf8()) {
        if (!buildRows.values) buildRows = rowNumbers(bts.numRows());
        cols.push_back(buildRows.c());
      } else {
        cols.push_back(bts.column(i).c());
      }
    }
  } catch (const std::exception& e) {
    return make_unexpected(e.what());
  }
  size_t nProbe = probeOut_.size();
  int64_t cap = n;
  int unique = 0;
  if (fpdb_hj_table_info(*table, nullptr, nullptr, nullptr, &unique) != FPDB_OK)
    return make_unexpected(fpdb_last_error());
  if (!unique && fpdb_hj_probe_count(ctx, *table, FPDB_JOIN_INNER, keys.data(), (int)keys.size(), n, &cap) != FPDB_OK)
    return make_unexpected(fpdb_last_error());
  const Schema& os = *outputSchema_;
  std::vector<Column> outCols;
  for (auto& f : os.fields)
    outCols.push_back(Column::allocate(ctx, f.type == FPDB_TYPE_UTF8 ? FPDB_TYPE_INT64 : f.type, cap, f.nullable));
  // fpdb_hj_probe writes probe columns, then build columns
  std::vector<fpdb_column> out;
  size_t nBuild = buildOut_.size();
  for (size_t j = 0; j < nProbe; ++j) out.push_back(outCols[nBuild + j].c());
  for (size_t j = 0; j < nBuild; ++j) out.push_back(outCols[j].c());
  int64_t produced = 0;
  fpdb_status s = fpdb_hj_probe(ctx, *table, FPDB_JOIN_INNER, keys.data(), (int)keys.size(), n, cols.data(),
                                (int)nProbe, cols.data() + nProbe, (int)nBuild, out.data(), cap, &produced);
  if (s != FPDB_OK) return make_unexpected(fpdb_last_error());
  if (produced == 0) return {};
  for (auto& c : outCols) c.length = produced;
  for (size_t j = 0; j < os.fields.size(); ++j) {
    if (os.fields[j].type != FPDB_TYPE_UTF8) continue;
    const Column& src = j < nBuild ? bts.column(buildOut_[j]) : probe->column(probeOut_[j - nBuild]);
    auto t = takeUtf8(src, outCols[j].c(), produced);
    if (!t) return make_unexpected(t.error());
    outCols[j] = *t;
  }
  return buffer(TupleSet::make(os, outCols));
}

// HashJoinProbeKernel::joinBuildTupleSetIndex (HashJoinProbeKernel.cpp:80-124):
// the new index against every probe row received so far.
Expected<void> HashJoinProbeKernel::joinBuildTupleSetIndex(const std::shared_ptr<TupleSetIndex>& tupleSetIndex) {
  bool first = !buildIndex_;
  auto r = putBuildTupleSetIndex(tupleSetIndex);
  if (!r) return r;
  if (!probeTupleSet_) return {};
  r = bufferOutputSchema(buildIndex_->getTupleSet()->schema(), probeTupleSet_->schema());
  if (!r) return r;
  if (joinType_ == LEFT_SEMI) return {};            // all semi rows come at finalize
  if (first) return joinPairs(*buildIndex_, probeTupleSet_);
  // the arrived rows only: a table of their own, dropped afterwards
  auto part = TupleSetIndex::make(tupleSetIndex->getColumnNames(), tupleSetIndex->getTupleSet());
  if (!part) return make_unexpected(part.error());
  return joinPairs(**part, probeTupleSet_);
}

// HashJoinProbeKernel::joinProbeTupleSet (HashJoinProbeKernel.cpp:126-168):
// the new probe rows against every build row received so far.
Expected<void> HashJoinProbeKernel::joinProbeTupleSet(const std::shared_ptr<TupleSet>& tupleSet) {
  auto r = putProbeTupleSet(tupleSet);
  if (!r) return r;
  if (!buildIndex_) return {};
  r = bufferOutputSchema(buildIndex_->getTupleSet()->schema(), probeTupleSet_->schema());
  if (!r) return r;
  if (joinType_ == LEFT_SEMI) return {};
  return joinPairs(*buildIndex_, tupleSet);
}

// Gathers rows of the listed columns of `in` into new columns typed as the
// output fields starting at firstField.
Expected<std::vector<Column>> HashJoinProbeKernel::gatherSide(const TupleSet& in, const std::vector<int>& cols,
                                                              const int64_t* d_rows, int64_t count,
                                                              size_t firstField) {
  fpdb_context* ctx = defaultContext();
  std::vector<Column> outCols;
  std::vector<fpdb_column> src, dst;
  fpdb_column rowCol{};
  rowCol.type = FPDB_TYPE_INT64;
  rowCol.length = count;
  rowCol.values = const_cast<int64_t*>(d_rows);
  for (size_t j = 0; j < cols.size(); ++j) {
    const Field& f = outputSchema_->fields[firstField + j];
    if (in.column(cols[j]).isUtf8()) {
      auto t = takeUtf8(in.column(cols[j]), rowCol, count);
      if (!t) return make_unexpected(t.error());
      outCols.push_back(*t);
      continue;
    }
    outCols.push_back(Column::allocate(ctx, f.type, count, f.nullable));
    src.push_back(in.column(cols[j]).c());
    dst.push_back(outCols.back().c());
  }
  if (!src.empty() && count > 0) {
    if (fpdb_gather(ctx, src.data(), (int)src.size(), d_rows, count, dst.data()) != FPDB_OK)
      return make_unexpected(fpdb_last_error());
  }
  for (auto& c : outCols) c.length = count;
  return outCols;
}

// Rows of `side` (a tracked table) that `other`'s keys never match, or match
// (matched = true), as device row indices.
static Expected<int64_t> sideRows(TupleSetIndex& side, const TupleSet* other, const std::vector<std::string>& otherKeys,
                                  bool matched, DeviceBuffer& rows) {
  fpdb_context* ctx = defaultContext();
  auto table = side.table(true, keyTypes(other->schema(), otherKeys));
  if (!table) return make_unexpected(table.error());
  if (other->numRows() > 0) {
    std::vector<fpdb_column> raw;
    for (auto& n : otherKeys) raw.push_back(other->column(other->schema().getFieldIndex(n)).c());
    Column keyIds;
    auto pk = side.probeKeys(raw, other->numRows(), &keyIds);
    if (!pk) return make_unexpected(pk.error());
    std::vector<fpdb_column> keys = *pk;
    int64_t produced = 0;
    if (fpdb_hj_probe(ctx, *table, FPDB_JOIN_LEFT_SEMI, keys.data(), (int)keys.size(), other->numRows(), nullptr, 0,
                      nullptr, 0, nullptr, 0, &produced) != FPDB_OK)
      return make_unexpected(fpdb_last_error());
  }
  int64_t count = 0;
  int64_t cap = std::max<int64_t>(side.size(), 1);
  if (fpdb_hj_build_side_rows(ctx, *table, matched ? 1 : 0, static_cast<int64_t*>(rows.data()), cap, &count) !=
      FPDB_OK)
    return make_unexpected(fpdb_last_error());
  return count;
}

// HashJoinProbeKernel::computeOuterJoin (HashJoinProbeKernel.cpp:198-226) and
// HashSemiJoinProbeKernel::finalize (HashSemiJoinProbeKernel.cpp:113-160).
// Which build rows and which probe rows ever matched is found by one more
// probe over all rows received: build rows with every probe key (LEFT, FULL,
// LEFT_SEMI), probe rows with every build key (RIGHT, FULL).
Expected<void> HashJoinProbeKernel::finalize() {
  if (!outputSchema_ || !buildIndex_ || !probeTupleSet_) return {};
  fpdb_context* ctx = defaultContext();
  const Schema& os = *outputSchema_;
  size_t nBuild = buildOut_.size();
  bool left = joinType_ == LEFT || joinType_ == FULL || joinType_ == LEFT_SEMI;
  bool right = joinType_ == RIGHT || joinType_ == FULL;
  if (left && buildIndex_->size() > 0) {
    DeviceBuffer rows(ctx, (uint64_t)buildIndex_->size() * 8);
    auto count = sideRows(*buildIndex_, probeTupleSet_.get(), pred_.rightColumnNames, joinType_ == LEFT_SEMI, rows);
    if (!count) return make_unexpected(count.error());
    auto cols = gatherSide(*buildIndex_->getTupleSet(), buildOut_, static_cast<int64_t*>(rows.data()), *count, 0);
    if (!cols) return make_unexpected(cols.error());
    for (size_t j = 0; j < probeOut_.size(); ++j) cols->push_back(Column::allNull(ctx, os.fields[nBuild + j].type, *count));
    checkStatus(fpdb_context_synchronize(ctx), "synchronize");
    if (*count > 0) {
      auto r = buffer(TupleSet::make(os, *cols));
      if (!r) return r;
    }
  }
  if (right && probeTupleSet_->numRows() > 0) {
    auto probeIndex = TupleSetIndex::make(pred_.rightColumnNames, probeTupleSet_);
    if (!probeIndex) return make_unexpected(probeIndex.error());
    DeviceBuffer rows(ctx, (uint64_t)probeTupleSet_->numRows() * 8);
    auto count = sideRows(**probeIndex, buildIndex_->getTupleSet().get(), buildIndex_->getColumnNames(), false, rows);
    if (!count) return make_unexpected(count.error());
    auto probeCols = gatherSide(*probeTupleSet_, probeOut_, static_cast<int64_t*>(rows.data()), *count, nBuild);
    if (!probeCols) return make_unexpected(probeCols.error());
    std::vector<Column> cols;
    for (size_t j = 0; j < nBuild; ++j) cols.push_back(Column::allNull(ctx, os.fields[j].type, *count));
    for (auto& c : *probeCols) cols.push_back(c);
    checkStatus(fpdb_context_synchronize(ctx), "synchronize");
    if (*count > 0) {
      auto r = buffer(TupleSet::make(os, cols));
      if (!r) return r;
    }
  }
  return {};
}

// HashJoinProbeAbstractKernel::buffer (HashJoinProbeAbstractKernel.cpp:48-58)
Expected<void> HashJoinProbeKernel::buffer(const std::shared_ptr<TupleSet>& tupleSet) {
  if (!buffer_) {
    buffer_ = std::make_shared<TupleSet>(tupleSet->schema(), tupleSet->columns());
    return {};
  }
  return (*buffer_)->append(tupleSet);
}

HashJoinProbePOp::HashJoinProbePOp(const std::string& name, const std::vector<std::string>& projectColumnNames,
                                   int nodeId, const HashJoinPredicate& pred, JoinType joinType)
    : PhysicalOp(name, POpType::HASH_JOIN_PROBE, projectColumnNames, nodeId),
      kernel_(pred, std::set<std::string>(projectColumnNames.begin(), projectColumnNames.end()), joinType) {
  // HashJoinProbePOp.cpp:15-48
  if (joinType == RIGHT_SEMI) throw std::runtime_error("Unsupported hash join type, RIGHT_SEMI");
}

// HashJoinProbePOp.cpp:54-117
void HashJoinProbePOp::onReceive(const Envelope& msg) {
  Expected<void> r;
  switch (msg.message().type()) {
    case MessageType::START: return;
    case MessageType::TUPLESET:
      r = kernel_.joinProbeTupleSet(static_cast<const TupleSetMessage&>(msg.message()).tuples());
      break;
    case MessageType::TUPLESET_INDEX:
      r = kernel_.joinBuildTupleSetIndex(static_cast<const TupleSetIndexMessage&>(msg.message()).getTupleSetIndex());
      break;
    case MessageType::COMPLETE:
      if (!ctx()->isComplete() && ctx()->operatorMap().allComplete(POpRelationshipType::Producer)) {
        r = kernel_.finalize();
        if (!r) {
          ctx()->notifyError(r.error() + ", " + name());
          return;
        }
        send(true);
        if (!sentResult_) sendEmpty();
        ctx()->notifyComplete();
      }
      return;
    default:
      ctx()->notifyError("Unrecognized message type " + msg.message().getTypeString() + ", " + name());
      return;
  }
  if (!r) {
    ctx()->notifyError(r.error() + ", " + name());
    return;
  }
  send(false);
}

void HashJoinProbePOp::send(bool force) {
  const auto& buffer = kernel_.getBuffer();
  if (!buffer) return;
  int64_t rows = (*buffer)->numRows();
  if (rows >= DefaultBufferSize || (force && rows > 0)) {
    ctx()->tell(std::make_shared<TupleSetMessage>(*buffer, name()));
    sentResult_ = true;
    kernel_.clearBuffer();
  }
}

// HashJoinProbePOp.cpp:132-140: an empty result still carries the output schema
void HashJoinProbePOp::sendEmpty() {
  const auto& schema = kernel_.getOutputSchema();
  if (!schema) {
    ctx()->notifyError("OutputSchema not set yet, " + name());
    return;
  }
  ctx()->tell(std::make_shared<TupleSetMessage>(TupleSet::make(*schema), name()));
  sentResult_ = true;
}

}  // namespace fpdb::amd
