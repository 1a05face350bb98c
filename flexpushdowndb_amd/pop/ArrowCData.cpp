#include "ArrowCData.h"

#include <cstring>
#include <memory>

namespace fpdb::amd {

const char* arrowFormat(fpdb_type t) {
  switch (t) {
    case FPDB_TYPE_INT8: return "c";
    case FPDB_TYPE_UINT8: return "C";
    case FPDB_TYPE_INT16: return "s";
    case FPDB_TYPE_UINT16: return "S";
    case FPDB_TYPE_INT32: return "i";
    case FPDB_TYPE_UINT32: return "I";
    case FPDB_TYPE_INT64: return "l";
    case FPDB_TYPE_UINT64: return "L";
    case FPDB_TYPE_FLOAT32: return "f";
    case FPDB_TYPE_FLOAT64: return "g";
    case FPDB_TYPE_DATE32: return "tdD";
    case FPDB_TYPE_DATE64: return "tdm";
    case FPDB_TYPE_UTF8: return "u";
  }
  return "";
}

Expected<fpdb_type> typeFromArrowFormat(const std::string& f) {
  static const std::pair<const char*, fpdb_type> table[] = {
      {"c", FPDB_TYPE_INT8},    {"C", FPDB_TYPE_UINT8},   {"s", FPDB_TYPE_INT16},   {"S", FPDB_TYPE_UINT16},
      {"i", FPDB_TYPE_INT32},   {"I", FPDB_TYPE_UINT32},  {"l", FPDB_TYPE_INT64},   {"L", FPDB_TYPE_UINT64},
      {"f", FPDB_TYPE_FLOAT32}, {"g", FPDB_TYPE_FLOAT64}, {"tdD", FPDB_TYPE_DATE32}, {"tdm", FPDB_TYPE_DATE64},
      {"u", FPDB_TYPE_UTF8},    {"U", FPDB_TYPE_UTF8}};
  for (auto& [name, t] : table)
    if (f == name) return t;
  return make_unexpected("Arrow format '" + f + "' is not a supported column type");
}

namespace {

// ---- export ------------------------------------------------------------------

struct SchemaPrivate {
  std::string format, name;
  std::vector<ArrowSchema> children;
  std::vector<ArrowSchema*> childPtrs;
};

void releaseSchema(ArrowSchema* s) {
  if (!s || !s->release) return;
  for (int64_t i = 0; i < s->n_children; ++i)
    if (s->children[i]->release) s->children[i]->release(s->children[i]);
  delete static_cast<SchemaPrivate*>(s->private_data);
  s->release = nullptr;
}

struct ArrayPrivate {
  std::vector<std::vector<uint8_t>> buffers;
  std::vector<const void*> bufferPtrs;
  std::vector<ArrowArray> children;
  std::vector<ArrowArray*> childPtrs;
};

void releaseArray(ArrowArray* a) {
  if (!a || !a->release) return;
  for (int64_t i = 0; i < a->n_children; ++i)
    if (a->children[i]->release) a->children[i]->release(a->children[i]);
  delete static_cast<ArrayPrivate*>(a->private_data);
  a->release = nullptr;
}

void fillSchema(ArrowSchema* s, SchemaPrivate* p, int64_t flags) {
  s->format = p->format.c_str();
  s->name = p->name.c_str();
  s->metadata = nullptr;
  s->flags = flags;
  s->n_children = (int64_t)p->children.size();
  p->childPtrs.clear();
  for (auto& c : p->children) p->childPtrs.push_back(&c);
  s->children = p->childPtrs.empty() ? nullptr : p->childPtrs.data();
  s->dictionary = nullptr;
  s->release = releaseSchema;
  s->private_data = p;
}

void fillArray(ArrowArray* a, ArrayPrivate* p, int64_t length, int64_t nullCount, bool hasValidity) {
  a->length = length;
  a->null_count = nullCount;
  a->offset = 0;
  p->bufferPtrs.clear();
  for (size_t i = 0; i < p->buffers.size(); ++i)
    p->bufferPtrs.push_back(i == 0 && !hasValidity ? nullptr : p->buffers[i].data());
  a->n_buffers = (int64_t)p->buffers.size();
  a->buffers = p->bufferPtrs.data();
  p->childPtrs.clear();
  for (auto& c : p->children) p->childPtrs.push_back(&c);
  a->n_children = (int64_t)p->children.size();
  a->children = p->childPtrs.empty() ? nullptr : p->childPtrs.data();
  a->dictionary = nullptr;
  a->release = releaseArray;
  a->private_data = p;
}

// ---- import ------------------------------------------------------------------

bool bitAt(const uint8_t* bits, int64_t i) { return (bits[i >> 3] >> (i & 7)) & 1; }

Expected<HostColumn> importColumn(const ArrowSchema* s, const ArrowArray* a, fpdb_type type, int64_t start,
                                  int64_t rows) {
  HostColumn h;
  h.type = type;
  h.length = rows;
  const bool large = std::string(s->format) == "U";
  const int64_t want = type == FPDB_TYPE_UTF8 ? 3 : 2;
  if (a->n_buffers != want)
    return make_unexpected("Arrow column '" + std::string(s->name ? s->name : "") + "' has " +
                           std::to_string(a->n_buffers) + " buffers, expected " + std::to_string(want));
  const uint8_t* validity = static_cast<const uint8_t*>(a->buffers[0]);
  if (validity && a->null_count != 0) {
    h.valid.resize((size_t)rows);
    for (int64_t r = 0; r < rows; ++r) h.valid[r] = bitAt(validity, start + r);
  }
  if (type == FPDB_TYPE_UTF8) {
    h.offsets.resize((size_t)rows + 1);
    int64_t base = 0;
    for (int64_t r = 0; r <= rows; ++r) {
      int64_t o = large ? static_cast<const int64_t*>(a->buffers[1])[start + r]
                        : (int64_t) static_cast<const int32_t*>(a->buffers[1])[start + r];
      if (r == 0) base = o;
      if (o - base > INT32_MAX) return make_unexpected("utf8 column holds more than 2^31 - 1 bytes");
      h.offsets[r] = (int32_t)(o - base);
    }
    const uint8_t* data = static_cast<const uint8_t*>(a->buffers[2]);
    h.bytes.assign(data ? data + base : nullptr, data ? data + base + h.offsets[rows] : nullptr);
  } else {
    const int w = typeWidth(type);
    const uint8_t* v = static_cast<const uint8_t*>(a->buffers[1]);
    h.bytes.resize((size_t)rows * w);
    if (rows) std::memcpy(h.bytes.data(), v + start * w, h.bytes.size());
  }
  return h;
}

}  // namespace

Expected<void> exportHostColumns(const Schema& schema, const std::vector<HostColumn>& columns, ArrowSchema* schemaOut,
                                 ArrowArray* arrayOut) {
  if (columns.size() != schema.fields.size()) return make_unexpected("export: column count differs from the schema");
  const int64_t n = columns.empty() ? 0 : columns[0].length;
  auto sp = std::make_unique<SchemaPrivate>();
  auto ap = std::make_unique<ArrayPrivate>();
  sp->format = "+s";
  sp->children.resize(columns.size());
  ap->buffers.resize(1);   // the struct's (absent) validity
  ap->children.resize(columns.size());
  for (auto& cs : sp->children) cs.release = nullptr;
  for (auto& ca : ap->children) ca.release = nullptr;
  // on an error, free the children filled so far
  struct Undo {
    SchemaPrivate* sp;
    ArrayPrivate* ap;
    ~Undo() {
      if (!sp) return;
      for (auto& cs : sp->children) releaseSchema(&cs);
      for (auto& ca : ap->children) releaseArray(&ca);
    }
  } undo{sp.get(), ap.get()};
  for (size_t c = 0; c < columns.size(); ++c) {
    const HostColumn& h = columns[c];
    if (h.length != n) return make_unexpected("export: columns differ in length");
    auto csp = new SchemaPrivate();
    csp->format = arrowFormat(schema.fields[c].type);
    csp->name = schema.fields[c].name;
    if (csp->format.empty()) {
      delete csp;
      return make_unexpected("export: unsupported column type");
    }
    fillSchema(&sp->children[c], csp, schema.fields[c].nullable || !h.valid.empty() ? ARROW_FLAG_NULLABLE : 0);
    auto cap = new ArrayPrivate();
    int64_t nulls = 0;
    std::vector<uint8_t> bits((size_t)(n + 7) / 8, 0);
    for (int64_t r = 0; r < n; ++r) {
      if (h.isValid(r))
        bits[r >> 3] |= (uint8_t)(1u << (r & 7));
      else
        ++nulls;
    }
    cap->buffers.push_back(std::move(bits));
    if (schema.fields[c].type == FPDB_TYPE_UTF8) {
      std::vector<uint8_t> offs((size_t)(n + 1) * 4);
      if (h.offsets.size() != (size_t)n + 1) {
        delete cap;
        return make_unexpected("export: utf8 column without offsets");
      }
      std::memcpy(offs.data(), h.offsets.data(), offs.size());
      cap->buffers.push_back(std::move(offs));
      cap->buffers.push_back(std::vector<uint8_t>(h.bytes.begin(), h.bytes.end()));
      if (cap->buffers.back().empty()) cap->buffers.back().push_back(0);   // a non-null data pointer
    } else {
      cap->buffers.push_back(std::vector<uint8_t>(h.bytes.begin(), h.bytes.end()));
      if (cap->buffers.back().empty()) cap->buffers.back().push_back(0);
    }
    fillArray(&ap->children[c], cap, n, nulls, nulls > 0);
  }
  undo.sp = nullptr;
  fillSchema(schemaOut, sp.release(), 0);
  fillArray(arrayOut, ap.release(), n, 0, false);
  return {};
}

Expected<void> importHostColumns(ArrowSchema* s, ArrowArray* a, Schema* schema, std::vector<HostColumn>* columns) {
  struct Releaser {
    ArrowSchema* s;
    ArrowArray* a;
    ~Releaser() {
      if (a && a->release) a->release(a);
      if (s && s->release) s->release(s);
    }
  } releaser{s, a};
  if (!s || !a || !s->release || !a->release) return make_unexpected("import: released or missing structures");
  if (std::string(s->format) != "+s") return make_unexpected("import: expected a struct array (format +s)");
  if (s->n_children != a->n_children) return make_unexpected("import: schema and array differ in children");
  schema->fields.clear();
  columns->clear();
  for (int64_t c = 0; c < s->n_children; ++c) {
    const ArrowSchema* cs = s->children[c];
    const ArrowArray* ca = a->children[c];
    if (cs->dictionary || ca->dictionary) return make_unexpected("import: dictionary-encoded columns are not supported");
    auto t = typeFromArrowFormat(cs->format);
    if (!t) return make_unexpected(t.error());
    const int64_t start = a->offset + ca->offset;
    if (ca->length < a->offset + a->length) return make_unexpected("import: a child is shorter than its struct");
    auto h = importColumn(cs, ca, *t, start, a->length);
    if (!h) return make_unexpected(h.error());
    schema->fields.push_back(Field{cs->name ? cs->name : "", *t, (cs->flags & ARROW_FLAG_NULLABLE) != 0});
    columns->push_back(std::move(*h));
  }
  return {};
}

Expected<void> exportTupleSet(const TupleSet& tupleSet, ArrowSchema* schemaOut, ArrowArray* arrayOut) {
  try {
    return exportHostColumns(tupleSet.schema(), tupleSet.toHost(), schemaOut, arrayOut);
  } catch (const std::exception& e) {
    return make_unexpected(e.what());
  }
}

Expected<std::shared_ptr<TupleSet>> importTupleSet(ArrowSchema* schemaIn, ArrowArray* arrayIn) {
  Schema schema;
  std::vector<HostColumn> cols;
  auto r = importHostColumns(schemaIn, arrayIn, &schema, &cols);
  if (!r) return make_unexpected(r.error());
  try {
    return TupleSet::fromHost(defaultContext(), schema, cols);
  } catch (const std::exception& e) {
    return make_unexpected(e.what());
  }
}

}  // namespace fpdb::amd
