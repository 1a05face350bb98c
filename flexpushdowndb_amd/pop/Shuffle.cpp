#include "Shuffle.h"

#include "Globals.h"
#include "KeyDictionary.h"

namespace fpdb::amd {

Expected<std::vector<std::shared_ptr<TupleSet>>> ShuffleKernel2::shuffle(
    const std::vector<std::string>& columnNames, size_t numSlots, const std::shared_ptr<TupleSet>& tupleSet) {
  // key columns, as RecordBatchHasher::make resolves them (RecordBatchHasher.cpp:27-35)
  std::vector<fpdb_column> keys;
  for (const auto& name : columnNames) {
    int i = tupleSet->schema().getFieldIndex(name);
    if (i < 0) return make_unexpected("Column '" + name + "' not found when creating RecordBatchHasher");
    if (tupleSet->column(i).isUtf8()) return make_unexpected("Shuffle on the utf8 column '" + name + "' is not supported");
    keys.push_back(tupleSet->column(i).c());
  }
  if (numSlots == 0) return make_unexpected("ShuffleKernel2: no slots");
  fpdb_context* ctx = defaultContext();
  const int64_t n = tupleSet->numRows();
  std::vector<fpdb_column> in, out;
  std::vector<Column> outCols;
  // utf8 columns travel as row numbers through the shuffle and are taken
  // afterwards in the shuffled order
  int rowsAt = -1;
  Column rows, rowsOut;
  try {
    for (int i = 0; i < tupleSet->numColumns(); ++i) {
      const Column& c = tupleSet->column(i);
      if (c.isUtf8()) {
        outCols.push_back(Column());
        continue;
      }
      in.push_back(c.c());
      outCols.push_back(Column::allocate(ctx, c.type, n, c.validity != nullptr));
      out.push_back(outCols.back().c());
    }
    for (auto& c : tupleSet->columns())
      if (c.isUtf8()) rowsAt = 0;
    if (rowsAt == 0) {
      rows = rowNumbers(n);
      rowsOut = Column::allocate(ctx, FPDB_TYPE_INT64, n, false);
      rowsAt = (int)in.size();
      in.push_back(rows.c());
      out.push_back(rowsOut.c());
    }
  } catch (const std::exception& e) {
    return make_unexpected(e.what());
  }
  std::vector<int64_t> offsets(numSlots + 1);
  if (fpdb_shuffle(ctx, keys.data(), (int)keys.size(), in.data(), (int)in.size(), n, (int)numSlots, out.data(),
                   nullptr, offsets.data()) != FPDB_OK)
    return make_unexpected(fpdb_last_error());
  if (rowsAt >= 0) {
    fpdb_column rc = rowsOut.c();
    for (int i = 0; i < tupleSet->numColumns(); ++i) {
      if (!tupleSet->column(i).isUtf8()) continue;
      auto t = takeUtf8(tupleSet->column(i), rc, n);
      if (!t) return make_unexpected(t.error());
      outCols[i] = *t;
    }
  }
  // slice the shuffled table into one TupleSet per slot (ShuffleKernel2.cpp:107-119)
  std::vector<std::shared_ptr<TupleSet>> slots;
  for (size_t p = 0; p < numSlots; ++p) {
    std::vector<Column> cols;
    for (auto& oc : outCols) cols.push_back(oc.slice(offsets[p], offsets[p + 1] - offsets[p]));
    slots.push_back(TupleSet::make(tupleSet->schema(), cols));
  }
  return slots;
}

ShufflePOp::ShufflePOp(const std::string& name, const std::vector<std::string>& projectColumnNames, int nodeId,
                       std::vector<std::string> shuffleColumnNames)
    : PhysicalOp(name, POpType::SHUFFLE, projectColumnNames, nodeId),
      shuffleColumnNames_(std::move(shuffleColumnNames)) {}

ShufflePOp::ShufflePOp(const std::string& name, const std::vector<std::string>& projectColumnNames, int nodeId,
                       std::vector<std::string> shuffleColumnNames, std::vector<std::string> consumerVec)
    : PhysicalOp(name, POpType::SHUFFLE, projectColumnNames, nodeId),
      shuffleColumnNames_(std::move(shuffleColumnNames)), consumerVec_(std::move(consumerVec)) {}

void ShufflePOp::produce(const std::shared_ptr<PhysicalOp>& op) {
  PhysicalOp::produce(op);
  consumerVec_.push_back(op->name());
}

void ShufflePOp::onReceive(const Envelope& msg) {
  const auto& m = msg.message();
  switch (m.type()) {
    case MessageType::START: onStart(); break;
    case MessageType::TUPLESET: onTupleSet(static_cast<const TupleSetMessage&>(m)); break;
    case MessageType::COMPLETE: onComplete(static_cast<const CompleteMessage&>(m)); break;
    default: ctx()->notifyError("Unrecognized message type: " + m.getTypeString() + ", " + name_);
  }
}

void ShufflePOp::onStart() { buffers_.assign(consumerVec_.size(), std::nullopt); }

// ShufflePOp.cpp:135-170: an empty input goes to every slot as it is.
void ShufflePOp::onTupleSet(const TupleSetMessage& msg) {
  const auto& tupleSet = msg.tuples();
  std::vector<std::shared_ptr<TupleSet>> shuffled;
  if (tupleSet->numRows() == 0) {
    shuffled.assign(consumerVec_.size(), tupleSet);
  } else {
    auto r = ShuffleKernel2::shuffle(shuffleColumnNames_, consumerVec_.size(), tupleSet);
    if (!r) {
      ctx()->notifyError(r.error() + ", " + name_);
      return;
    }
    shuffled = std::move(*r);
  }
  for (int p = 0; p < (int)shuffled.size(); ++p) {
    auto b = buffer(shuffled[p], p);
    if (b) b = send(p, false);
    if (!b) ctx()->notifyError(b.error());
  }
}

void ShufflePOp::onComplete(const CompleteMessage&) {
  if (!ctx()->isComplete() && ctx()->operatorMap().allComplete(POpRelationshipType::Producer)) {
    for (int p = 0; p < (int)buffers_.size(); ++p) {
      auto s = send(p, true);
      if (!s) ctx()->notifyError(s.error());
    }
    ctx()->notifyComplete();
  }
}

// ShufflePOp.cpp:91-111
Expected<void> ShufflePOp::buffer(const std::shared_ptr<TupleSet>& tupleSet, int partitionIndex) {
  auto& slot = buffers_[partitionIndex];
  if (!slot.has_value()) {
    slot = tupleSet;
    return {};
  }
  auto r = TupleSet::concatenate({*slot, tupleSet});
  if (!r) return make_unexpected(r.error());
  slot = *r;
  return {};
}

// ShufflePOp.cpp:113-133 (the compute-side branch: the buffer goes to its consumer)
Expected<void> ShufflePOp::send(int partitionIndex, bool force) {
  auto& slot = buffers_[partitionIndex];
  if (slot.has_value() && (force || (*slot)->numRows() >= DefaultBufferSize)) {
    ctx()->tell(std::make_shared<TupleSetMessage>(*slot, name_), {consumerVec_[partitionIndex]});
    slot.reset();
  }
  return {};
}

}  // namespace fpdb::amd
