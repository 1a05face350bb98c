// Key encoding for the join and group operators: key columns the device
// kernels do not take as they are (utf8, more than two columns, two wider
// than 4 bytes, nullable group keys) go through a device key dictionary that
// gives each distinct key tuple a dense int32 id (fpdb_keydict_*).  The role of
// Arrow's key encoders inside acero's hash join and Grouper
// (HashJoinArrowKernel.cpp:138-145, GroupArrowKernel.cpp:245-291).
#pragma once

#include <vector>

#include "TupleSet.h"

namespace fpdb::amd {

// Whether the hash-join kernels take these key columns directly: one integer
// or date column, or two integer/date columns of at most 4 bytes each.
bool joinKeysNative(const std::vector<fpdb_type>& types);

class KeyDictionary {
 public:
  explicit KeyDictionary(const std::vector<fpdb_type>& types, int hashBits = 64);
  ~KeyDictionary();
  KeyDictionary(const KeyDictionary&) = delete;
  KeyDictionary& operator=(const KeyDictionary&) = delete;

  // INT32 ids of n rows with a validity bitmap: insert adds new tuples;
  // without insert an absent tuple is null.  nullIsKey makes null a key
  // value (group-by); otherwise a row with a null key column is null (join).
  Expected<Column> encode(const std::vector<fpdb_column>& keys, int64_t n, bool insert, bool nullIsKey);
  // Values of key column `key` for ids (null ids give nulls).
  Expected<Column> decode(int key, const Column& ids, int64_t n);
  int64_t size() const;
  const std::vector<fpdb_type>& types() const { return types_; }

 private:
  fpdb_keydict* d_ = nullptr;
  std::vector<fpdb_type> types_;
};

// Rows of a utf8 column named by an INT64/INT32 row-number column (null or
// negative: a null row), as a new nullable utf8 column.
Expected<Column> takeUtf8(const Column& src, const fpdb_column& rows, int64_t n);

// An INT64 column 0, 1, ..., n - 1 (row numbers carried through a kernel that
// moves fixed-width columns only, to resolve utf8 columns afterwards).
Column rowNumbers(int64_t n);

}  // namespace fpdb::amd
