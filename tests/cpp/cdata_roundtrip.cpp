// ctypes hooks for tests/test_arrow_c_data.py: a struct array comes in through
// the Arrow C Data Interface (exported by pyarrow), is imported into host
// columns -- or into a device TupleSet -- and goes back out the same way.
#include <cstdio>

#include "ArrowCData.h"

using namespace fpdb::amd;

extern "C" int cdata_host_roundtrip(ArrowSchema* in_s, ArrowArray* in_a, ArrowSchema* out_s, ArrowArray* out_a,
                                    char* err, int err_len) {
  Schema schema;
  std::vector<HostColumn> cols;
  auto r = importHostColumns(in_s, in_a, &schema, &cols);
  if (r) r = exportHostColumns(schema, cols, out_s, out_a);
  if (!r) {
    std::snprintf(err, (size_t)err_len, "%s", r.error().c_str());
    return 1;
  }
  return 0;
}

extern "C" int cdata_device_roundtrip(ArrowSchema* in_s, ArrowArray* in_a, ArrowSchema* out_s, ArrowArray* out_a,
                                      char* err, int err_len) {
  auto ts = importTupleSet(in_s, in_a);
  Expected<void> r;
  if (ts)
    r = exportTupleSet(**ts, out_s, out_a);
  else
    r = make_unexpected(ts.error());
  if (!r) {
    std::snprintf(err, (size_t)err_len, "%s", r.error().c_str());
    return 1;
  }
  return 0;
}
