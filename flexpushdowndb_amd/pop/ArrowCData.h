// The Arrow C Data Interface at the TupleSet boundary: a TupleSet leaves as an
// ArrowSchema + ArrowArray pair (a struct array, one child per column -- the
// form arrow::ExportRecordBatch writes and arrow::ImportRecordBatch reads) and
// comes back the same way, so the reference's arrow::Table-based TupleSet
// (fpdb-tuple/include/fpdb/tuple/TupleSet.h:31-192) converts to and from this
// build's device TupleSet without linking libarrow.  The two structs are the
// ABI the Arrow specification fixes ("The Arrow C data interface",
// arrow/c/abi.h); their layout has to be exactly this.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "TupleSet.h"

#ifndef ARROW_C_DATA_INTERFACE
#define ARROW_C_DATA_INTERFACE

#define ARROW_FLAG_DICTIONARY_ORDERED 1
#define ARROW_FLAG_NULLABLE 2
#define ARROW_FLAG_MAP_KEYS_SORTED 4

extern "C" {
struct ArrowSchema {
  const char* format;
  const char* name;
  const char* metadata;
  int64_t flags;
  int64_t n_children;
  struct ArrowSchema** children;
  struct ArrowSchema* dictionary;
  void (*release)(struct ArrowSchema*);
  void* private_data;
};

struct ArrowArray {
  int64_t length;
  int64_t null_count;
  int64_t offset;
  int64_t n_buffers;
  int64_t n_children;
  const void** buffers;
  struct ArrowArray** children;
  struct ArrowArray* dictionary;
  void (*release)(struct ArrowArray*);
  void* private_data;
};
}

#endif  // ARROW_C_DATA_INTERFACE

namespace fpdb::amd {

// Arrow format string of a column type ("l", "i", "u", "tdD", ...), and back
// (an unsupported format gives an error naming it).
const char* arrowFormat(fpdb_type t);
Expected<fpdb_type> typeFromArrowFormat(const std::string& format);

// Host columns <-> a struct ArrowArray.  Export fills *schema and *array with
// structures the consumer owns and frees through their release callbacks
// (the buffers are copies).  Import copies what the arrays hold -- honouring
// each array's offset, null count and missing validity buffer -- and then
// releases the imported structures, as a consumer that moved them must.
Expected<void> exportHostColumns(const Schema& schema, const std::vector<HostColumn>& columns, ArrowSchema* schemaOut,
                                 ArrowArray* arrayOut);
Expected<void> importHostColumns(ArrowSchema* schemaIn, ArrowArray* arrayIn, Schema* schema,
                                 std::vector<HostColumn>* columns);

// The same for device TupleSets (a copy through host memory each way).
Expected<void> exportTupleSet(const TupleSet& tupleSet, ArrowSchema* schemaOut, ArrowArray* arrayOut);
Expected<std::shared_ptr<TupleSet>> importTupleSet(ArrowSchema* schemaIn, ArrowArray* arrayIn);

}  // namespace fpdb::amd
