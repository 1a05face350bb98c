#include "KeyDictionary.h"

#include <stdexcept>

namespace fpdb::amd {

bool joinKeysNative(const std::vector<fpdb_type>& types) {
  if (types.empty() || types.size() > 2) return false;
  for (auto t : types) {
    if (t == FPDB_TYPE_UTF8 || t == FPDB_TYPE_FLOAT32 || t == FPDB_TYPE_FLOAT64) return false;
    if (types.size() == 2 && typeWidth(t) > 4) return false;
  }
  return true;
}

KeyDictionary::KeyDictionary(const std::vector<fpdb_type>& types, int hashBits) : types_(types) {
  std::vector<int32_t> t(types.begin(), types.end());
  checkStatus(fpdb_keydict_create(defaultContext(), t.data(), (int)t.size(), hashBits, &d_), "fpdb_keydict_create");
}

KeyDictionary::~KeyDictionary() {
  if (d_) fpdb_keydict_destroy(d_);
}

int64_t KeyDictionary::size() const {
  int64_t n = 0;
  fpdb_keydict_size(d_, &n);
  return n;
}

Expected<Column> KeyDictionary::encode(const std::vector<fpdb_column>& keys, int64_t n, bool insert,
                                       bool nullIsKey) {
  try {
    Column ids = Column::allocate(defaultContext(), FPDB_TYPE_INT32, n, true);
    fpdb_column o = ids.c();
    if (fpdb_keydict_encode(defaultContext(), d_, keys.data(), (int)keys.size(), n, insert ? 1 : 0,
                            nullIsKey ? 1 : 0, &o, nullptr) != FPDB_OK)
      return make_unexpected(fpdb_last_error());
    return ids;
  } catch (const std::exception& e) {
    return make_unexpected(e.what());
  }
}

Expected<Column> KeyDictionary::decode(int key, const Column& ids, int64_t n) {
  fpdb_context* ctx = defaultContext();
  fpdb_column ic = ids.c();
  try {
    Column out;
    if (types_[key] == FPDB_TYPE_UTF8) {
      int64_t bytes = 0;
      if (fpdb_keydict_decode_bytes(ctx, d_, key, &ic, n, &bytes) != FPDB_OK) return make_unexpected(fpdb_last_error());
      out = Column::allocateUtf8(ctx, n, bytes, true);
    } else {
      out = Column::allocate(ctx, types_[key], n, true);
    }
    fpdb_column oc = out.c();
    if (fpdb_keydict_decode(ctx, d_, key, &ic, n, &oc) != FPDB_OK) return make_unexpected(fpdb_last_error());
    return out;
  } catch (const std::exception& e) {
    return make_unexpected(e.what());
  }
}

Expected<Column> takeUtf8(const Column& src, const fpdb_column& rows, int64_t n) {
  fpdb_context* ctx = defaultContext();
  fpdb_column in = src.c();
  int64_t bytes = 0;
  if (fpdb_utf8_take_bytes(ctx, &in, &rows, n, &bytes) != FPDB_OK) return make_unexpected(fpdb_last_error());
  try {
    Column out = Column::allocateUtf8(ctx, n, bytes, true);
    fpdb_column oc = out.c();
    if (fpdb_utf8_take(ctx, &in, &rows, n, &oc) != FPDB_OK) return make_unexpected(fpdb_last_error());
    return out;
  } catch (const std::exception& e) {
    return make_unexpected(e.what());
  }
}

Column rowNumbers(int64_t n) {
  Column c = Column::allocate(defaultContext(), FPDB_TYPE_INT64, n, false);
  checkStatus(fpdb_iota_int64(defaultContext(), static_cast<int64_t*>(c.values->data()), n, 0), "fpdb_iota_int64");
  return c;
}

}  // namespace fpdb::amd
