// TupleSetIndex (fpdb-tuple/include/fpdb/tuple/TupleSetIndex.h:25-110): the rows
// of a hash join build side plus the names of its key columns.  The reference
// keeps a host multimap from key to row; here the rows stay in HBM and the
// device hash table over the keys (fpdb_hj_build) is built when a probe first
// needs it and dropped whenever rows are added.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "KeyDictionary.h"
#include "TupleSet.h"

namespace fpdb::amd {

class TupleSetIndex {
 public:
  TupleSetIndex(std::vector<std::string> columnNames, std::shared_ptr<TupleSet> tupleSet);
  ~TupleSetIndex();
  TupleSetIndex(const TupleSetIndex&) = delete;
  TupleSetIndex& operator=(const TupleSetIndex&) = delete;

  // TupleSetIndex::make (fpdb-tuple/src/TupleSetIndex.cpp:21-41): every key
  // column must exist in the tuple set.
  static Expected<std::shared_ptr<TupleSetIndex>> make(const std::vector<std::string>& columnNames,
                                                       const std::shared_ptr<TupleSet>& tupleSet);

  int64_t size() const { return tupleSet_->numRows(); }
  const std::shared_ptr<TupleSet>& getTupleSet() const { return tupleSet_; }
  const std::vector<std::string>& getColumnNames() const { return columnNames_; }

  // Appends rows whose schema holds the same fields, in any order
  // (TupleSetIndex::put, TupleSetIndex.cpp:88-128).
  Expected<void> put(const std::shared_ptr<TupleSet>& tupleSet);
  // Appends another index's rows after this one's (TupleSetIndex::merge,
  // TupleSetIndex.cpp:130-166): the key column names must be equal.
  Expected<void> merge(const std::shared_ptr<TupleSetIndex>& other);

  // The device hash table over the key columns; trackMatches asks for one
  // matched flag per row (fpdb_hj_build's track_build_matches).  A table built
  // with tracking is returned for either request; asking for tracking after a
  // table without it was built rebuilds it.  probeTypes are the key types of
  // the side that will probe it: when either side's keys are not ones the
  // join kernels take directly, the table is built over key-dictionary ids
  // and probes go through probeKeys.  Owned by the index.
  Expected<fpdb_hash_table*> table(bool trackMatches, const std::vector<fpdb_type>& probeTypes);
  // Probe key columns as the table takes them: the columns themselves, or
  // their dictionary ids (held in *holder, which must outlive the probe).
  Expected<std::vector<fpdb_column>> probeKeys(const std::vector<fpdb_column>& keys, int64_t n, Column* holder);
  // Drops the device table (its matched flags included).
  void dropTable();

 private:
  std::vector<std::string> columnNames_;
  std::shared_ptr<TupleSet> tupleSet_;
  fpdb_hash_table* table_ = nullptr;
  bool tableTracks_ = false;
  std::unique_ptr<KeyDictionary> dict_;
  Column keyIds_;
};

}  // namespace fpdb::amd
