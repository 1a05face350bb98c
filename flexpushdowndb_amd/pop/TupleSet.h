// Device-resident TupleSet: the in/out contract of the reference's physical
// operators (fpdb-tuple/include/fpdb/tuple/TupleSet.h:31-192), with Arrow
// column layout (values + LSB-first validity bitmap) kept in HBM.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "Expected.h"
#include "fpdb_amd.h"

namespace fpdb::amd {

// Owns one device allocation made through an fpdb_context.
class DeviceBuffer {
 public:
  DeviceBuffer(fpdb_context* ctx, uint64_t bytes);
  // Borrows memory owned elsewhere (a cached segment); `release` runs instead
  // of the free when the last holder lets go.
  DeviceBuffer(fpdb_context* ctx, void* ptr, uint64_t bytes, std::function<void()> release);
  ~DeviceBuffer();
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  void* data() const { return ptr_; }
  uint64_t size() const { return bytes_; }
  fpdb_context* ctx() const { return ctx_; }

 private:
  fpdb_context* ctx_;
  void* ptr_ = nullptr;
  uint64_t bytes_;
  std::function<void()> release_;
};

// Device context shared by the operators of one executor.
fpdb_context* defaultContext();
// Throws std::runtime_error with the library's message on a failed status.
void checkStatus(fpdb_status s, const char* what);

int typeWidth(fpdb_type t);
std::string typeName(fpdb_type t);

struct Field {
  std::string name;
  fpdb_type type;
  bool nullable = false;
};

struct Schema {
  std::vector<Field> fields;
  int getFieldIndex(const std::string& name) const;
  std::string toString() const;
};

// One column of a TupleSet: a single contiguous device chunk, or a slice of
// one (rows [offset, offset + length) of shared buffers, like arrow::Array::Slice).
// A utf8 column keeps Arrow's layout: `values` holds int32 offsets (one more
// than its rows), `data` the characters.
struct Column {
  fpdb_type type = FPDB_TYPE_INT64;
  int64_t length = 0;
  int64_t offset = 0;
  std::shared_ptr<DeviceBuffer> values;
  std::shared_ptr<DeviceBuffer> validity;   // null: no nulls
  std::shared_ptr<DeviceBuffer> data;       // utf8 characters
  fpdb_column c() const;
  bool isUtf8() const { return type == FPDB_TYPE_UTF8; }
  // Bytes of characters in rows [offset, offset + length) (reads two offsets
  // back from the device); 0 for other types.
  int64_t utf8Bytes() const;

  static Column allocate(fpdb_context* ctx, fpdb_type type, int64_t length, bool nullable);
  // A utf8 column with room for `dataBytes` characters.
  static Column allocateUtf8(fpdb_context* ctx, int64_t length, int64_t dataBytes, bool nullable);
  // Room for `length` rows taken from `src` without repeats (a compaction or a
  // reordering): same type and nullability, and for utf8 src's characters.
  static Column allocateLike(fpdb_context* ctx, const Column& src, int64_t length);
  static Column allNull(fpdb_context* ctx, fpdb_type type, int64_t length);
  // Rows [begin, begin + rows) of this column, sharing its buffers.
  Column slice(int64_t begin, int64_t rows) const;
};

// Host copy of a column, for tests and display.  utf8: `bytes` holds the
// characters and `offsets` (length + 1 entries, from 0) the row boundaries.
struct HostColumn {
  fpdb_type type;
  std::vector<uint8_t> bytes;   // length * width (utf8: the characters)
  std::vector<bool> valid;      // empty: all valid
  std::vector<int32_t> offsets; // utf8 only
  int64_t length = 0;
  int64_t int64At(int64_t i) const;
  double doubleAt(int64_t i) const;
  std::string stringAt(int64_t i) const;
  bool isValid(int64_t i) const { return valid.empty() || valid[i]; }
  // A utf8 column of the given strings (nullopt-free: use `valid` for nulls).
  static HostColumn utf8(const std::vector<std::string>& strings);
};

class TupleSet {
 public:
  TupleSet(Schema schema, std::vector<Column> columns);

  static std::shared_ptr<TupleSet> make(const Schema& schema);
  static std::shared_ptr<TupleSet> make(const Schema& schema, std::vector<Column> columns);
  // Uploads host columns (values as raw little-endian bytes).
  static std::shared_ptr<TupleSet> fromHost(fpdb_context* ctx, const Schema& schema,
                                            const std::vector<HostColumn>& columns);
  static Expected<std::shared_ptr<TupleSet>> concatenate(const std::vector<std::shared_ptr<TupleSet>>& sets);

  int64_t numRows() const { return numRows_; }
  int numColumns() const { return (int)columns_.size(); }
  const Schema& schema() const { return schema_; }
  const std::vector<Column>& columns() const { return columns_; }
  const Column& column(int i) const { return columns_[i]; }
  Expected<Column> getColumnByName(const std::string& name) const;

  // Appends another TupleSet with the same schema (a device copy).
  Expected<void> append(const std::shared_ptr<TupleSet>& other);
  // Keeps the named columns that exist, in the given order.
  Expected<std::shared_ptr<TupleSet>> projectExist(const std::vector<std::string>& names) const;

  std::vector<HostColumn> toHost() const;
  std::string showString(int64_t maxRows = 20) const;

 private:
  Schema schema_;
  std::vector<Column> columns_;
  int64_t numRows_ = 0;
};

}  // namespace fpdb::amd
