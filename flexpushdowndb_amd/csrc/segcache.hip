// Device-resident segment cache: column segments of table partitions kept in
// HBM under (partition, column) keys with a size-bounded LRU replacement
// policy (SegmentCache, fpdb-cache/src/SegmentCache.cpp:21-49;
// LRUCachingPolicy, fpdb-cache/src/policy/LRUCachingPolicy.cpp:18-73).
//
// Host code only: the policy and the key map stay on the host, the segment
// data lives in HBM (a store copies the column in once; a load hands out the
// device buffers without copying).  A loaded segment is pinned until
// released, so an eviction never frees memory a query still reads: an
// evicted, pinned segment leaves the map and the policy at once (its size is
// free for the policy, as in the reference) and its memory is returned when
// the last pin goes.
#include <list>
#include <map>
#include <memory>

#include "fpdb_internal.hpp"

namespace fpdb {

struct CachedSegment {
  std::string partition, column;
  int64_t start = 0, finish = 0;
  int32_t type = 0;
  int64_t length = 0;
  void* values = nullptr;
  uint8_t* validity = nullptr;
  uint64_t size = 0;        // values + validity bytes (Arrow buffer sizes)
  int pins = 0;
  bool in_cache = true;
};

}  // namespace fpdb

using namespace fpdb;

struct fpdb_segment_cache {
  Context* ctx = nullptr;
  uint64_t capacity = 0;
  uint64_t free_size = 0;
  // usage queue, most recently used first (LRUCachingPolicy::usageQueue_)
  std::list<std::shared_ptr<CachedSegment>> lru;
  // (partition, column) -> position in the usage queue.  The key leaves the
  // range out, as SegmentKey::operator== and SegmentKey::hash do
  // (fpdb-cache/src/SegmentKey.cpp:58-72).
  std::map<std::pair<std::string, std::string>, std::list<std::shared_ptr<CachedSegment>>::iterator> index;
  // pinned segments, by values pointer (evicted ones included)
  std::map<const void*, std::shared_ptr<CachedSegment>> pinned;
  int64_t hits = 0, misses = 0;
};

namespace fpdb {

static void free_segment(Context* c, CachedSegment& s) {
  if (s.values) ctx_free(c, s.values);
  if (s.validity) ctx_free(c, s.validity);
  s.values = nullptr;
  s.validity = nullptr;
}

// Takes a segment out of the map and the usage queue; its memory goes now
// unless it is pinned.
static void evict(fpdb_segment_cache* k, std::list<std::shared_ptr<CachedSegment>>::iterator it) {
  std::shared_ptr<CachedSegment> s = *it;
  k->index.erase({s->partition, s->column});
  k->lru.erase(it);
  k->free_size += s->size;
  s->in_cache = false;
  if (s->pins == 0) free_segment(k->ctx, *s);
}

static bool is_device_pointer(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice;
}

}  // namespace fpdb

extern "C" fpdb_status fpdb_cache_create(fpdb_context* ctx, uint64_t capacity_bytes, int policy,
                                         fpdb_segment_cache** out) {
  FPDB_CHECK(ctx && out, FPDB_ERR_INVALID, "fpdb_cache_create: NULL argument");
  FPDB_CHECK(policy == FPDB_CACHE_LRU, FPDB_ERR_INVALID, "fpdb_cache_create: unknown caching policy");
  auto* k = new fpdb_segment_cache();
  k->ctx = as_ctx(ctx);
  k->capacity = capacity_bytes;
  k->free_size = capacity_bytes;
  *out = k;
  return FPDB_OK;
}

extern "C" fpdb_status fpdb_cache_destroy(fpdb_segment_cache* k) {
  if (!k) return FPDB_OK;
  for (auto& s : k->lru) free_segment(k->ctx, *s);
  for (auto& kv : k->pinned) free_segment(k->ctx, *kv.second);
  delete k;
  return FPDB_OK;
}

extern "C" fpdb_status fpdb_cache_store(fpdb_segment_cache* k, const char* partition, const char* column,
                                        int64_t start_offset, int64_t finish_offset, const fpdb_column* data,
                                        int* stored) {
  FPDB_CHECK(k && partition && column && data && stored, FPDB_ERR_INVALID, "fpdb_cache_store: NULL argument");
  *stored = 0;
  const int w = type_width(data->type);
  FPDB_CHECK(w > 0, FPDB_ERR_INVALID, "fpdb_cache_store: unsupported column type");
  FPDB_CHECK(data->length >= 0 && data->offset >= 0, FPDB_ERR_INVALID, "fpdb_cache_store: bad length or offset");
  FPDB_CHECK(data->length == 0 || data->values, FPDB_ERR_INVALID, "fpdb_cache_store: NULL values");
  const uint64_t vbytes = (uint64_t)data->length * w;
  const uint64_t bbytes = data->validity ? (uint64_t)(data->length + 7) / 8 : 0;
  const uint64_t size = vbytes + bbytes;
  const auto key = std::make_pair(std::string(partition), std::string(column));
  // a key already present is refreshed, not stored twice
  auto found = k->index.find(key);
  if (found != k->index.end()) {
    k->lru.splice(k->lru.begin(), k->lru, found->second);
    *stored = 1;
    return FPDB_OK;
  }
  // LRUCachingPolicy::onStore: too large for the whole cache -> not cached;
  // otherwise evict from the back until it fits
  if (size > k->capacity) return FPDB_OK;
  while (k->free_size < size) evict(k, std::prev(k->lru.end()));

  auto s = std::make_shared<CachedSegment>();
  s->partition = key.first;
  s->column = key.second;
  s->start = start_offset;
  s->finish = finish_offset;
  s->type = data->type;
  s->length = data->length;
  s->size = size;
  Context* c = k->ctx;
  FPDB_TRY(ctx_alloc(c, vbytes, &s->values));
  if (bbytes) {
    fpdb_status st = ctx_alloc(c, (bbytes + 3) / 4 * 4, reinterpret_cast<void**>(&s->validity));
    if (st != FPDB_OK) {
      free_segment(c, *s);
      return st;
    }
  }
  hipError_t e = hipSuccess;
  if (vbytes)
    e = hipMemcpyAsync(s->values, static_cast<const uint8_t*>(data->values) + data->offset * w, vbytes,
                       hipMemcpyDefault, c->stream);
  fpdb_status st = FPDB_OK;
  if (e == hipSuccess && bbytes) {
    if ((data->offset & 7) == 0) {
      e = hipMemcpyAsync(s->validity, data->validity + data->offset / 8, bbytes, hipMemcpyDefault, c->stream);
    } else if (is_device_pointer(data->validity)) {
      // bit-shifted copy on the device
      fpdb_column dst{data->type, 0, data->length, 0, s->values, s->validity};
      st = launch_fill(c, s->validity, 0, (bbytes + 3) / 4 * 4);
      if (st == FPDB_OK) st = fpdb_copy_rows(reinterpret_cast<fpdb_context*>(c), data, 1, data->length, &dst, 0);
    } else {
      st = fail(FPDB_ERR_INVALID, "fpdb_cache_store: a host validity bitmap must start on a byte boundary");
    }
  }
  if (e != hipSuccess) st = fail(FPDB_ERR_HIP, std::string("fpdb_cache_store: copy failed: ") + hipGetErrorString(e));
  if (st != FPDB_OK) {
    free_segment(c, *s);
    return st;
  }
  k->lru.push_front(s);
  k->index[key] = k->lru.begin();
  k->free_size -= size;
  *stored = 1;
  return FPDB_OK;
}

extern "C" fpdb_status fpdb_cache_load(fpdb_segment_cache* k, const char* partition, const char* column,
                                       int64_t start_offset, int64_t finish_offset, fpdb_column* out, int* hit) {
  FPDB_CHECK(k && partition && column && out && hit, FPDB_ERR_INVALID, "fpdb_cache_load: NULL argument");
  (void)start_offset;
  (void)finish_offset;
  *hit = 0;
  auto found = k->index.find({partition, column});
  if (found == k->index.end()) {
    ++k->misses;
    return FPDB_OK;
  }
  // LRUCachingPolicy::onLoad: move to the front of the usage queue
  k->lru.splice(k->lru.begin(), k->lru, found->second);
  std::shared_ptr<CachedSegment> s = *found->second;
  ++s->pins;
  k->pinned[s->values] = s;
  ++k->hits;
  *out = fpdb_column{s->type, 0, s->length, 0, s->values, s->validity};
  *hit = 1;
  return FPDB_OK;
}

extern "C" fpdb_status fpdb_cache_release(fpdb_segment_cache* k, const void* values) {
  FPDB_CHECK(k, FPDB_ERR_INVALID, "fpdb_cache_release: NULL cache");
  auto it = k->pinned.find(values);
  FPDB_CHECK(it != k->pinned.end(), FPDB_ERR_INVALID, "fpdb_cache_release: segment is not pinned");
  std::shared_ptr<CachedSegment> s = it->second;
  if (--s->pins == 0) {
    k->pinned.erase(it);
    if (!s->in_cache) free_segment(k->ctx, *s);
  }
  return FPDB_OK;
}

extern "C" fpdb_status fpdb_cache_remove(fpdb_segment_cache* k, const char* partition, const char* column,
                                         int64_t* removed) {
  FPDB_CHECK(k && partition, FPDB_ERR_INVALID, "fpdb_cache_remove: NULL argument");
  int64_t n = 0;
  for (auto it = k->lru.begin(); it != k->lru.end();) {
    auto next = std::next(it);
    if ((*it)->partition == partition && (column == nullptr || (*it)->column == column)) {
      evict(k, it);
      ++n;
    }
    it = next;
  }
  if (removed) *removed = n;
  return FPDB_OK;
}

extern "C" fpdb_status fpdb_cache_clear(fpdb_segment_cache* k) {
  FPDB_CHECK(k, FPDB_ERR_INVALID, "fpdb_cache_clear: NULL cache");
  while (!k->lru.empty()) evict(k, k->lru.begin());
  k->free_size = k->capacity;
  k->hits = k->misses = 0;
  return FPDB_OK;
}

extern "C" fpdb_status fpdb_cache_info(fpdb_segment_cache* k, int64_t* num_segments, uint64_t* used_bytes,
                                       uint64_t* free_bytes, int64_t* hits, int64_t* misses) {
  FPDB_CHECK(k, FPDB_ERR_INVALID, "fpdb_cache_info: NULL cache");
  if (num_segments) *num_segments = (int64_t)k->lru.size();
  if (used_bytes) *used_bytes = k->capacity - k->free_size;
  if (free_bytes) *free_bytes = k->free_size;
  if (hits) *hits = k->hits;
  if (misses) *misses = k->misses;
  return FPDB_OK;
}

extern "C" fpdb_status fpdb_cache_layout(fpdb_segment_cache* k, char* buf, uint64_t buf_len, uint64_t* needed) {
  FPDB_CHECK(k, FPDB_ERR_INVALID, "fpdb_cache_layout: NULL cache");
  std::string s;
  for (auto& seg : k->lru) s += seg->partition + "\t" + seg->column + "\n";
  if (needed) *needed = s.size() + 1;
  if (buf && buf_len) {
    const uint64_t n = std::min<uint64_t>(buf_len - 1, s.size());
    std::memcpy(buf, s.data(), n);
    buf[n] = 0;
  }
  return FPDB_OK;
}
