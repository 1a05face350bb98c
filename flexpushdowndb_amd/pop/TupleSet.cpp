#include "TupleSet.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <sstream>
#include <stdexcept>

namespace fpdb::amd {

void checkStatus(fpdb_status s, const char* what) {
  if (s != FPDB_OK) throw std::runtime_error(std::string(what) + ": " + fpdb_last_error());
}

fpdb_context* defaultContext() {
  static std::once_flag once;
  static fpdb_context* ctx = nullptr;
  std::call_once(once, [] { checkStatus(fpdb_context_create(0, 0, nullptr, &ctx), "fpdb_context_create"); });
  return ctx;
}

DeviceBuffer::DeviceBuffer(fpdb_context* ctx, uint64_t bytes) : ctx_(ctx), bytes_(bytes) {
  checkStatus(fpdb_device_alloc(ctx, bytes ? bytes : 1, &ptr_), "fpdb_device_alloc");
}

DeviceBuffer::DeviceBuffer(fpdb_context* ctx, void* ptr, uint64_t bytes, std::function<void()> release)
    : ctx_(ctx), ptr_(ptr), bytes_(bytes), release_(std::move(release)) {}

DeviceBuffer::~DeviceBuffer() {
  if (release_)
    release_();
  else if (ptr_)
    fpdb_device_free(ctx_, ptr_);
}

int typeWidth(fpdb_type t) {
  switch (t) {
    case FPDB_TYPE_INT8: case FPDB_TYPE_UINT8: return 1;
    case FPDB_TYPE_INT16: case FPDB_TYPE_UINT16: return 2;
    case FPDB_TYPE_INT32: case FPDB_TYPE_UINT32: case FPDB_TYPE_FLOAT32: case FPDB_TYPE_DATE32: return 4;
    default: return 8;
  }
}

std::string typeName(fpdb_type t) {
  switch (t) {
    case FPDB_TYPE_INT8: return "int8";
    case FPDB_TYPE_INT16: return "int16";
    case FPDB_TYPE_INT32: return "int32";
    case FPDB_TYPE_INT64: return "int64";
    case FPDB_TYPE_UINT8: return "uint8";
    case FPDB_TYPE_UINT16: return "uint16";
    case FPDB_TYPE_UINT32: return "uint32";
    case FPDB_TYPE_UINT64: return "uint64";
    case FPDB_TYPE_FLOAT32: return "float";
    case FPDB_TYPE_FLOAT64: return "double";
    case FPDB_TYPE_DATE32: return "date32";
    case FPDB_TYPE_DATE64: return "date64";
    case FPDB_TYPE_UTF8: return "utf8";
  }
  return "?";
}

int Schema::getFieldIndex(const std::string& name) const {
  for (size_t i = 0; i < fields.size(); ++i)
    if (fields[i].name == name) return (int)i;
  return -1;
}

std::string Schema::toString() const {
  std::ostringstream o;
  for (size_t i = 0; i < fields.size(); ++i)
    o << (i ? ", " : "") << fields[i].name << ": " << typeName(fields[i].type);
  return o.str();
}

static uint64_t validityBytes(int64_t n) { return (uint64_t)((n + 31) / 32) * 4; }

fpdb_column Column::c() const {
  fpdb_column col;
  col.type = type;
  col.reserved = 0;
  col.length = length;
  col.offset = offset;
  col.values = values ? values->data() : nullptr;
  col.validity = validity ? static_cast<uint8_t*>(validity->data()) : nullptr;
  col.data = data ? data->data() : nullptr;
  col.data_bytes = data ? (int64_t)data->size() : 0;
  return col;
}

int64_t Column::utf8Bytes() const {
  if (!isUtf8() || length == 0) return 0;
  int32_t ends[2];
  const uint8_t* offs = static_cast<const uint8_t*>(values->data());
  checkStatus(fpdb_copy_to_host(defaultContext(), &ends[0], offs + offset * 4, 4), "copy");
  checkStatus(fpdb_copy_to_host(defaultContext(), &ends[1], offs + (offset + length) * 4, 4), "copy");
  return (int64_t)ends[1] - ends[0];
}

Column Column::slice(int64_t begin, int64_t rows) const {
  Column s = *this;
  s.offset = offset + begin;
  s.length = rows;
  return s;
}

Column Column::allocateUtf8(fpdb_context* ctx, int64_t length, int64_t dataBytes, bool nullable) {
  Column c;
  c.type = FPDB_TYPE_UTF8;
  c.length = length;
  c.values = std::make_shared<DeviceBuffer>(ctx, (uint64_t)(length + 1) * 4);
  checkStatus(fpdb_fill(ctx, c.values->data(), 0, c.values->size()), "fpdb_fill");
  c.data = std::make_shared<DeviceBuffer>(ctx, (uint64_t)std::max<int64_t>(dataBytes, 1));
  if (nullable) {
    c.validity = std::make_shared<DeviceBuffer>(ctx, validityBytes(length));
    checkStatus(fpdb_fill(ctx, c.validity->data(), 0, c.validity->size()), "fpdb_fill");
  }
  return c;
}

Column Column::allocateLike(fpdb_context* ctx, const Column& src, int64_t length) {
  if (src.isUtf8()) return allocateUtf8(ctx, length, src.utf8Bytes(), src.validity != nullptr);
  return allocate(ctx, src.type, length, src.validity != nullptr);
}

Column Column::allocate(fpdb_context* ctx, fpdb_type type, int64_t length, bool nullable) {
  if (type == FPDB_TYPE_UTF8) return allocateUtf8(ctx, length, 0, nullable);
  Column c;
  c.type = type;
  c.length = length;
  c.values = std::make_shared<DeviceBuffer>(ctx, (uint64_t)length * typeWidth(type));
  if (nullable) {
    c.validity = std::make_shared<DeviceBuffer>(ctx, validityBytes(length));
    checkStatus(fpdb_fill(ctx, c.validity->data(), 0, c.validity->size()), "fpdb_fill");
  }
  return c;
}

Column Column::allNull(fpdb_context* ctx, fpdb_type type, int64_t length) {
  Column c = allocate(ctx, type, length, true);
  checkStatus(fpdb_fill(ctx, c.values->data(), 0, c.values->size()), "fpdb_fill");
  return c;
}

HostColumn HostColumn::utf8(const std::vector<std::string>& strings) {
  HostColumn h;
  h.type = FPDB_TYPE_UTF8;
  h.length = (int64_t)strings.size();
  h.offsets.push_back(0);
  for (auto& s : strings) {
    h.bytes.insert(h.bytes.end(), s.begin(), s.end());
    h.offsets.push_back((int32_t)h.bytes.size());
  }
  return h;
}

std::string HostColumn::stringAt(int64_t i) const {
  if (type != FPDB_TYPE_UTF8) return std::to_string(int64At(i));
  return std::string(reinterpret_cast<const char*>(bytes.data()) + offsets[i], (size_t)(offsets[i + 1] - offsets[i]));
}

int64_t HostColumn::int64At(int64_t i) const {
  switch (type) {
    case FPDB_TYPE_INT8: return (int8_t)bytes[i];
    case FPDB_TYPE_UINT8: return bytes[i];
    case FPDB_TYPE_INT16: { int16_t v; std::memcpy(&v, &bytes[i * 2], 2); return v; }
    case FPDB_TYPE_UINT16: { uint16_t v; std::memcpy(&v, &bytes[i * 2], 2); return v; }
    case FPDB_TYPE_INT32: case FPDB_TYPE_DATE32: { int32_t v; std::memcpy(&v, &bytes[i * 4], 4); return v; }
    case FPDB_TYPE_UINT32: { uint32_t v; std::memcpy(&v, &bytes[i * 4], 4); return v; }
    case FPDB_TYPE_FLOAT32: { float v; std::memcpy(&v, &bytes[i * 4], 4); return (int64_t)v; }
    case FPDB_TYPE_FLOAT64: { double v; std::memcpy(&v, &bytes[i * 8], 8); return (int64_t)v; }
    default: { int64_t v; std::memcpy(&v, &bytes[i * 8], 8); return v; }
  }
}

double HostColumn::doubleAt(int64_t i) const {
  if (type == FPDB_TYPE_FLOAT64) { double v; std::memcpy(&v, &bytes[i * 8], 8); return v; }
  if (type == FPDB_TYPE_FLOAT32) { float v; std::memcpy(&v, &bytes[i * 4], 4); return v; }
  return (double)int64At(i);
}

TupleSet::TupleSet(Schema schema, std::vector<Column> columns)
    : schema_(std::move(schema)), columns_(std::move(columns)) {
  numRows_ = columns_.empty() ? 0 : columns_[0].length;
}

// Zero-row columns still own (tiny) device buffers, so every column handed to
// the C ABI has a values pointer.
std::shared_ptr<TupleSet> TupleSet::make(const Schema& schema) {
  std::vector<Column> cols;
  for (auto& f : schema.fields) cols.push_back(Column::allocate(defaultContext(), f.type, 0, f.nullable));
  return std::make_shared<TupleSet>(schema, cols);
}

std::shared_ptr<TupleSet> TupleSet::make(const Schema& schema, std::vector<Column> columns) {
  if (columns.size() != schema.fields.size())
    throw std::runtime_error("TupleSet::make: column count differs from the schema");
  return std::make_shared<TupleSet>(schema, std::move(columns));
}

std::shared_ptr<TupleSet> TupleSet::fromHost(fpdb_context* ctx, const Schema& schema,
                                             const std::vector<HostColumn>& hcols) {
  std::vector<Column> cols;
  for (size_t i = 0; i < hcols.size(); ++i) {
    const HostColumn& h = hcols[i];
    bool nullable = !h.valid.empty();
    Column c;
    if (schema.fields[i].type == FPDB_TYPE_UTF8) {
      if ((int64_t)h.offsets.size() != h.length + 1) throw std::runtime_error("TupleSet::fromHost: utf8 offsets");
      c = Column::allocateUtf8(ctx, h.length, (int64_t)h.bytes.size(), nullable);
      std::vector<int32_t> offs(h.offsets.begin(), h.offsets.end());
      for (auto& o : offs) o -= h.offsets[0];
      checkStatus(fpdb_copy_to_device(ctx, c.values->data(), offs.data(), offs.size() * 4), "copy");
      if (!h.bytes.empty())
        checkStatus(fpdb_copy_to_device(ctx, c.data->data(), h.bytes.data() + h.offsets[0],
                                        (uint64_t)(h.offsets.back() - h.offsets[0])),
                    "copy");
    } else {
      c = Column::allocate(ctx, schema.fields[i].type, h.length, nullable);
      checkStatus(fpdb_copy_to_device(ctx, c.values->data(), h.bytes.data(), h.bytes.size()), "copy");
    }
    if (nullable) {
      std::vector<uint8_t> bits(validityBytes(h.length), 0);
      for (int64_t r = 0; r < h.length; ++r)
        if (h.valid[r]) bits[r >> 3] |= (uint8_t)(1u << (r & 7));
      checkStatus(fpdb_copy_to_device(ctx, c.validity->data(), bits.data(), bits.size()), "copy");
    }
    cols.push_back(c);
  }
  return make(schema, cols);
}

Expected<Column> TupleSet::getColumnByName(const std::string& name) const {
  int i = schema_.getFieldIndex(name);
  if (i < 0) return make_unexpected("Column '" + name + "' does not exist");
  return columns_[i];
}

// Device concatenation of two columns (values and validity bits).
static Column concatColumns(fpdb_context* ctx, const Column& a, const Column& b) {
  int64_t n = a.length + b.length;
  bool nullable = a.validity || b.validity;
  Column out = a.isUtf8() ? Column::allocateUtf8(ctx, n, a.utf8Bytes() + b.utf8Bytes(), nullable)
                          : Column::allocate(ctx, a.type, n, nullable);
  fpdb_column o = out.c();
  fpdb_column ia = a.c(), ib = b.c();
  checkStatus(fpdb_copy_rows(ctx, &ia, 1, a.length, &o, 0), "fpdb_copy_rows");
  checkStatus(fpdb_copy_rows(ctx, &ib, 1, b.length, &o, a.length), "fpdb_copy_rows");
  return out;
}

Expected<void> TupleSet::append(const std::shared_ptr<TupleSet>& other) {
  if (other->numColumns() != numColumns())
    return make_unexpected("TupleSet::append: column count differs");
  if (other->numRows() == 0) return {};
  if (numRows_ == 0) {
    columns_ = other->columns_;
    numRows_ = other->numRows_;
    return {};
  }
  fpdb_context* ctx = defaultContext();
  for (int i = 0; i < numColumns(); ++i) {
    if (columns_[i].type != other->columns_[i].type)
      return make_unexpected("TupleSet::append: column type differs for '" + schema_.fields[i].name + "'");
    try {
      columns_[i] = concatColumns(ctx, columns_[i], other->columns_[i]);
    } catch (const std::exception& e) {
      return make_unexpected(e.what());
    }
  }
  numRows_ += other->numRows_;
  return {};
}

Expected<std::shared_ptr<TupleSet>> TupleSet::concatenate(const std::vector<std::shared_ptr<TupleSet>>& sets) {
  if (sets.empty()) return make_unexpected("TupleSet::concatenate: no inputs");
  auto out = std::make_shared<TupleSet>(sets[0]->schema(), sets[0]->columns());
  for (size_t i = 1; i < sets.size(); ++i) {
    auto r = out->append(sets[i]);
    if (!r) return make_unexpected(r.error());
  }
  return out;
}

Expected<std::shared_ptr<TupleSet>> TupleSet::projectExist(const std::vector<std::string>& names) const {
  Schema s;
  std::vector<Column> cols;
  for (auto& n : names) {
    int i = schema_.getFieldIndex(n);
    if (i < 0) continue;
    s.fields.push_back(schema_.fields[i]);
    cols.push_back(columns_[i]);
  }
  return std::make_shared<TupleSet>(s, cols);
}

std::vector<HostColumn> TupleSet::toHost() const {
  fpdb_context* ctx = defaultContext();
  std::vector<HostColumn> out;
  for (auto& c : columns_) {
    HostColumn h;
    h.type = c.type;
    h.length = c.length;
    if (c.isUtf8()) {
      h.offsets.resize((size_t)c.length + 1, 0);
      checkStatus(fpdb_copy_to_host(ctx, h.offsets.data(), static_cast<uint8_t*>(c.values->data()) + c.offset * 4,
                                    h.offsets.size() * 4),
                  "copy");
      const int32_t base = h.offsets[0];
      for (auto& o : h.offsets) o -= base;
      h.bytes.resize((size_t)h.offsets.back());
      if (!h.bytes.empty())
        checkStatus(fpdb_copy_to_host(ctx, h.bytes.data(), static_cast<uint8_t*>(c.data->data()) + base,
                                      h.bytes.size()),
                    "copy");
    } else {
      h.bytes.resize((size_t)c.length * typeWidth(c.type));
      const int w = typeWidth(c.type);
      if (c.length)
        checkStatus(fpdb_copy_to_host(ctx, h.bytes.data(), static_cast<uint8_t*>(c.values->data()) + c.offset * w,
                                      h.bytes.size()),
                    "copy");
    }
    if (c.validity && c.length) {
      const int64_t b0 = c.offset >> 3, sh = c.offset & 7;
      std::vector<uint8_t> bits((sh + c.length + 7) / 8);
      checkStatus(fpdb_copy_to_host(ctx, bits.data(), static_cast<uint8_t*>(c.validity->data()) + b0, bits.size()),
                  "copy");
      h.valid.resize(c.length);
      for (int64_t r = 0; r < c.length; ++r) h.valid[r] = (bits[(r + sh) >> 3] >> ((r + sh) & 7)) & 1;
    }
    out.push_back(std::move(h));
  }
  return out;
}

std::string TupleSet::showString(int64_t maxRows) const {
  std::ostringstream o;
  o << schema_.toString() << "\n";
  auto host = toHost();
  for (int64_t r = 0; r < std::min<int64_t>(numRows_, maxRows); ++r) {
    for (size_t c = 0; c < host.size(); ++c) {
      o << (c ? " | " : "");
      if (!host[c].isValid(r)) o << "null";
      else if (host[c].type == FPDB_TYPE_UTF8) o << host[c].stringAt(r);
      else if (host[c].type == FPDB_TYPE_FLOAT64 || host[c].type == FPDB_TYPE_FLOAT32) o << host[c].doubleAt(r);
      else o << host[c].int64At(r);
    }
    o << "\n";
  }
  return o.str();
}

}  // namespace fpdb::amd
