#include "TupleSetIndex.h"

namespace fpdb::amd {

TupleSetIndex::TupleSetIndex(std::vector<std::string> columnNames, std::shared_ptr<TupleSet> tupleSet)
    : columnNames_(std::move(columnNames)), tupleSet_(std::move(tupleSet)) {}

TupleSetIndex::~TupleSetIndex() { dropTable(); }

void TupleSetIndex::dropTable() {
  if (table_) fpdb_hj_table_destroy(table_);
  table_ = nullptr;
  tableTracks_ = false;
  dict_.reset();
  keyIds_ = Column();
}

Expected<std::shared_ptr<TupleSetIndex>> TupleSetIndex::make(const std::vector<std::string>& columnNames,
                                                             const std::shared_ptr<TupleSet>& tupleSet) {
  if (columnNames.empty()) return make_unexpected("Cannot make TupleSetIndex. No key columns");
  for (auto& n : columnNames)
    if (tupleSet->schema().getFieldIndex(n) < 0)
      return make_unexpected("Cannot make TupleSetIndex. Column '" + n + "' does not exist");
  // a private column list, so later appends never change the sender's tuple set
  return std::make_shared<TupleSetIndex>(columnNames,
                                         std::make_shared<TupleSet>(tupleSet->schema(), tupleSet->columns()));
}

Expected<void> TupleSetIndex::put(const std::shared_ptr<TupleSet>& tupleSet) {
  const Schema& mine = tupleSet_->schema();
  const Schema& theirs = tupleSet->schema();
  if (theirs.fields.size() != mine.fields.size()) return make_unexpected("Schemas not compatible");
  std::vector<Column> aligned;
  for (auto& f : mine.fields) {
    int i = theirs.getFieldIndex(f.name);
    if (i < 0 || theirs.fields[i].type != f.type) return make_unexpected("Schemas not compatible");
    aligned.push_back(tupleSet->column(i));
  }
  if (tupleSet->numRows() == 0) return {};
  dropTable();
  return tupleSet_->append(TupleSet::make(mine, aligned));
}

Expected<void> TupleSetIndex::merge(const std::shared_ptr<TupleSetIndex>& other) {
  if (columnNames_ != other->columnNames_) return make_unexpected("Hash column names are incompatible");
  return put(other->tupleSet_);
}

Expected<fpdb_hash_table*> TupleSetIndex::table(bool trackMatches, const std::vector<fpdb_type>& probeTypes) {
  if (table_ && (tableTracks_ || !trackMatches)) return table_;
  dropTable();
  std::vector<fpdb_column> keys;
  std::vector<fpdb_type> types;
  for (auto& n : columnNames_) {
    const int i = tupleSet_->schema().getFieldIndex(n);
    keys.push_back(tupleSet_->column(i).c());
    types.push_back(tupleSet_->schema().fields[i].type);
  }
  if (!joinKeysNative(types) || !joinKeysNative(probeTypes)) {
    try {
      dict_ = std::make_unique<KeyDictionary>(types);
    } catch (const std::exception& e) {
      return make_unexpected(e.what());
    }
    auto ids = dict_->encode(keys, size(), true, false);
    if (!ids) return make_unexpected(ids.error());
    keyIds_ = *ids;
    keys = {keyIds_.c()};
  }
  fpdb_status s = fpdb_hj_build(defaultContext(), keys.data(), (int)keys.size(), size(), nullptr,
                                trackMatches ? 1 : 0, &table_);
  if (s != FPDB_OK) {
    table_ = nullptr;
    return make_unexpected(fpdb_last_error());
  }
  tableTracks_ = trackMatches;
  return table_;
}

Expected<std::vector<fpdb_column>> TupleSetIndex::probeKeys(const std::vector<fpdb_column>& keys, int64_t n,
                                                            Column* holder) {
  if (!dict_) return keys;
  auto ids = dict_->encode(keys, n, false, false);
  if (!ids) return make_unexpected(ids.error());
  *holder = *ids;
  return std::vector<fpdb_column>{holder->c()};
}

}  // namespace fpdb::amd
