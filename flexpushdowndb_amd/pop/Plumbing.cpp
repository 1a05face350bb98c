#include "Plumbing.h"

namespace fpdb::amd {

void SourcePOp::onReceive(const Envelope& msg) {
  if (msg.message().type() != MessageType::START) return;
  for (auto& b : batches_) ctx()->tell(std::make_shared<TupleSetMessage>(b, name_));
  ctx()->notifyComplete();
}

Expected<std::shared_ptr<TupleSet>> FilterPOp::filter(const std::shared_ptr<TupleSet>& ts, const std::string& column,
                                                      fpdb_cmp_op op, int64_t intConstant, double floatConstant,
                                                      const std::string* stringConstant) {
  int ci = ts->schema().getFieldIndex(column);
  if (ci < 0) return make_unexpected("Column '" + column + "' does not exist");
  fpdb_context* ctx = defaultContext();
  int64_t n = ts->numRows();
  if (n == 0) return ts;
  DeviceBuffer bitmap(ctx, (uint64_t)((n + 63) / 64) * 8);
  fpdb_column c = ts->column(ci).c();
  fpdb_status st;
  if (ts->column(ci).isUtf8()) {
    if (!stringConstant) return make_unexpected("Column '" + column + "' is utf8: compare it with a string");
    st = fpdb_compare_bitmap_utf8(ctx, &c, op, stringConstant->data(), (int64_t)stringConstant->size(), nullptr, n,
                                  static_cast<uint8_t*>(bitmap.data()), nullptr);
  } else {
    if (stringConstant) return make_unexpected("Column '" + column + "' is not utf8: compare it with a number");
    st = fpdb_compare_bitmap(ctx, &c, op, intConstant, floatConstant, nullptr, n,
                             static_cast<uint8_t*>(bitmap.data()), nullptr);
  }
  if (st != FPDB_OK) return make_unexpected(fpdb_last_error());
  std::vector<fpdb_column> in, out;
  std::vector<Column> outCols;
  for (int i = 0; i < ts->numColumns(); ++i) {
    in.push_back(ts->column(i).c());
    outCols.push_back(Column::allocateLike(ctx, ts->column(i), n));
    out.push_back(outCols.back().c());
  }
  int64_t kept = 0;
  if (fpdb_filter_compact(ctx, in.data(), (int)in.size(), n, static_cast<uint8_t*>(bitmap.data()), out.data(),
                          &kept) != FPDB_OK)
    return make_unexpected(fpdb_last_error());
  for (auto& oc : outCols) oc.length = kept;
  return TupleSet::make(ts->schema(), outCols);
}

void FilterPOp::onReceive(const Envelope& msg) {
  const auto& m = msg.message();
  if (m.type() == MessageType::TUPLESET) {
    auto ts = static_cast<const TupleSetMessage&>(m).tuples();
    auto r = filter(ts, column_, op_, intConstant_, floatConstant_, hasString_ ? &stringConstant_ : nullptr);
    if (!r) {
      ctx()->notifyError(r.error());
      return;
    }
    auto projected = projectColumnNames_.empty() ? *r : (*r)->projectExist(projectColumnNames_).value();
    ctx()->tell(std::make_shared<TupleSetMessage>(projected, name_));
  } else if (m.type() == MessageType::COMPLETE) {
    if (!ctx()->isComplete() && ctx()->operatorMap().allComplete(POpRelationshipType::Producer))
      ctx()->notifyComplete();
  }
}

void CollatePOp::onReceive(const Envelope& msg) {
  const auto& m = msg.message();
  if (m.type() == MessageType::TUPLESET) {
    received_.push_back(static_cast<const TupleSetMessage&>(m).tuples());
  } else if (m.type() == MessageType::COMPLETE) {
    if (!ctx()->isComplete() && ctx()->operatorMap().allComplete(POpRelationshipType::Producer))
      ctx()->notifyComplete();
  }
}

Expected<std::shared_ptr<TupleSet>> CollatePOp::tupleSet() const {
  if (received_.empty()) return make_unexpected("no TupleSet received");
  return TupleSet::concatenate(received_);
}

}  // namespace fpdb::amd
