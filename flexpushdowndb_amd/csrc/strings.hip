// utf8 columns and composite keys.
//
//   * take: rows of a utf8 column gathered by row number (lengths, a scan into
//     int32 offsets, a byte copy) -- joins, compaction and concatenation carry
//     string columns with it;
//   * comparison of a utf8 column with a constant string into a selection
//     bitmap;
//   * the key dictionary: each distinct tuple of key columns (any mix of
//     integer, float, date and utf8 columns, nulls included) gets a dense int32
//     id, so the join and group-by kernels, which take one integer key, run on
//     string, wide and nullable keys.
//
// Key dictionary.  An open-addressing table of 64-bit tuple hashes; each slot
// also holds the id of its tuple and, while a batch is being inserted, the
// smallest row that claimed it.  The distinct tuples are stored once in the
// dictionary's own columns (in id order), and every lookup compares the row's
// values with the stored tuple, so equality is exact whatever the hashes do.
// Inserting a batch runs in passes, each over a list of rows (all rows first):
//   insert  -- each row walks the probe sequence of its hash: a slot of the
//              same hash with an id compares the tuple (equal: that id);
//              an empty slot is claimed by compare-and-swap; a slot of the
//              same hash claimed in this pass is joined; either way the row
//              lowers the slot's representative row with atomicMin;
//   verify  -- after the kernel boundary the representative of every claimed
//              slot is final: a row equal to it takes the slot, a row that is
//              not (a genuine hash collision) goes on the next pass's list;
//   assign  -- representatives in row order get the next ids and their tuples
//              are appended to the stored columns; then every row of the pass
//              reads its slot's id.
// No thread ever waits for another (a wave cannot wait on one of its own
// lanes), and each pass settles at least the smallest row of every claimed
// slot, so the passes end; with full 64-bit hashes a second pass needs two
// distinct tuples with one hash.
#include <algorithm>

#include "compact.hpp"
#include "fpdb_internal.hpp"

namespace fpdb {
namespace {

constexpr int kStrThreads = 256;
constexpr int kDictMaxKeys = 8;
constexpr uint64_t kEmptySlot = ~0ULL;     // slot never claimed; also the hash of a row with no id
constexpr int64_t kMaxConstLen = 256;

// ---- views ---------------------------------------------------------------------

struct Utf8In {
  const int32_t* offs;
  const uint8_t* data;
  const uint8_t* validity;
  int64_t offset;
};

// A column of row numbers (INT64 or INT32, optional validity).  A null or
// negative entry names no row; one at or past `limit` raises the error flag
// and names no row.
struct RowsIn {
  const uint8_t* values;
  const uint8_t* validity;
  int64_t offset;
  int32_t wide;
  int64_t limit;
};

__device__ __forceinline__ int64_t rows_at(const RowsIn& r, int64_t i, int* err) {
  const int64_t k = i + r.offset;
  if (!valid_bit(r.validity, k)) return -1;
  const int64_t v = r.wide ? reinterpret_cast<const int64_t*>(r.values)[k]
                           : (int64_t) reinterpret_cast<const int32_t*>(r.values)[k];
  if (v < 0) return -1;
  if (v >= r.limit) {
    *err = 1;
    return -1;
  }
  return v;
}

// Length of the string at row `row` (0 for null or no row); `valid` tells
// whether the output row is non-null.
__device__ __forceinline__ int32_t span_of(const Utf8In& u, int64_t row, int32_t* start, bool* valid) {
  if (row < 0) {
    *valid = false;
    *start = 0;
    return 0;
  }
  const int64_t k = row + u.offset;
  if (!valid_bit(u.validity, k)) {
    *valid = false;
    *start = 0;
    return 0;
  }
  *valid = true;
  *start = u.offs[k];
  return u.offs[k + 1] - *start;
}

// Exclusive prefix sum over the threads of a workgroup of W waves; *total
// receives the workgroup's sum.  Every thread must call it.
template <int W>
__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t* total) {
  __shared__ int64_t s_w[W];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int64_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int64_t y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) s_w[w] = x;
  __syncthreads();
  int64_t base = 0, t = 0;
#pragma unroll
  for (int i = 0; i < W; ++i) {
    base += i < w ? s_w[i] : 0;
    t += s_w[i];
  }
  __syncthreads();
  *total = t;
  return base + x - v;
}

// ---- take -----------------------------------------------------------------------

// Rows per thread so that the workgroups fit the partial-sum array.
inline int take_per_thread(int64_t n) {
  return (int)std::max<int64_t>(1, ceil_div(n, (int64_t)kMaxPartials * kStrThreads));
}

__global__ __launch_bounds__(kStrThreads) void take_len_kernel(Utf8In in, RowsIn rows, int64_t n, int per,
                                                                unsigned long long* partials, int* err) {
  const int64_t r0 = ((int64_t)blockIdx.x * kStrThreads + threadIdx.x) * per;
  int64_t mine = 0;
  for (int j = 0; j < per; ++j) {
    const int64_t i = r0 + j;
    if (i >= n) break;
    int32_t s;
    bool v;
    mine += span_of(in, rows_at(rows, i, err), &s, &v);
  }
  int64_t total;
  (void)block_excl_scan<kStrThreads / 64>(mine, &total);
  if (threadIdx.x == 0) partials[blockIdx.x] = (unsigned long long)total;
}

// One workgroup: exclusive scan of `n` partial sums in place; the total goes
// to *total.
__global__ __launch_bounds__(1024) void scan_partials_kernel(unsigned long long* partials, int n, int64_t* total) {
  constexpr int kPer = kMaxPartials / 1024;
  const int i0 = threadIdx.x * kPer;
  int64_t v[kPer];
  int64_t mine = 0;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    v[j] = i0 + j < n ? (int64_t)partials[i0 + j] : 0;
    mine += v[j];
  }
  int64_t t;
  int64_t run = block_excl_scan<16>(mine, &t);
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    if (i0 + j < n) partials[i0 + j] = (unsigned long long)run;
    run += v[j];
  }
  if (threadIdx.x == 0) *total = t;
}

__global__ __launch_bounds__(kStrThreads) void take_offsets_kernel(Utf8In in, RowsIn rows, int64_t n, int per,
                                                                    const unsigned long long* bases,
                                                                    int32_t* out_offs, uint8_t* out_valid,
                                                                    int64_t out_row, int64_t base_bytes, int* err) {
  const int64_t r0 = ((int64_t)blockIdx.x * kStrThreads + threadIdx.x) * per;
  int64_t mine = 0;
  for (int j = 0; j < per; ++j) {
    const int64_t i = r0 + j;
    if (i >= n) break;
    int32_t s;
    bool v;
    mine += span_of(in, rows_at(rows, i, err), &s, &v);
  }
  int64_t total;
  int64_t pos = base_bytes + (int64_t)bases[blockIdx.x] + block_excl_scan<kStrThreads / 64>(mine, &total);
  if (blockIdx.x == 0 && threadIdx.x == 0) out_offs[out_row] = (int32_t)base_bytes;
  for (int j = 0; j < per; ++j) {
    const int64_t i = r0 + j;
    if (i >= n) break;
    int32_t s;
    bool v;
    pos += span_of(in, rows_at(rows, i, err), &s, &v);
    out_offs[out_row + i + 1] = (int32_t)pos;
    if (out_valid && v) {
      const int64_t b = out_row + i;
      atomicOr(reinterpret_cast<unsigned int*>(out_valid) + (b >> 5), 1u << (b & 31));
    }
  }
}

__global__ __launch_bounds__(kStrThreads) void take_copy_kernel(Utf8In in, RowsIn rows, int64_t n,
                                                                 const int32_t* out_offs, uint8_t* out_data,
                                                                 int64_t out_row, int* err) {
  for (int64_t i = (int64_t)blockIdx.x * kStrThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kStrThreads) {
    int32_t s;
    bool v;
    const int32_t len = span_of(in, rows_at(rows, i, err), &s, &v);
    if (len <= 0) continue;
    const uint8_t* src = in.data + s;
    uint8_t* dst = out_data + out_offs[out_row + i];
    for (int32_t b = 0; b < len; ++b) dst[b] = src[b];
  }
}

fpdb_status utf8_view(const fpdb_column* col, int64_t num_rows, Utf8In* u, const char* what) {
  FPDB_CHECK(col, FPDB_ERR_INVALID, std::string(what) + ": NULL column");
  FPDB_CHECK(col->type == FPDB_TYPE_UTF8, FPDB_ERR_INVALID, std::string(what) + ": column is not utf8");
  FPDB_CHECK(col->offset >= 0, FPDB_ERR_INVALID, std::string(what) + ": negative offset");
  FPDB_CHECK(num_rows <= 0 || col->values, FPDB_ERR_INVALID, std::string(what) + ": NULL offsets");
  FPDB_CHECK(num_rows <= 0 || col->data || col->data_bytes == 0, FPDB_ERR_INVALID,
             std::string(what) + ": NULL character data");
  u->offs = static_cast<const int32_t*>(col->values);
  u->data = static_cast<const uint8_t*>(col->data);
  u->validity = col->validity;
  u->offset = col->offset;
  return FPDB_OK;
}

fpdb_status rows_view(const fpdb_column* rows, int64_t num_rows, int64_t limit, RowsIn* r, const char* what) {
  FPDB_CHECK(rows, FPDB_ERR_INVALID, std::string(what) + ": NULL row numbers");
  FPDB_CHECK(rows->type == FPDB_TYPE_INT64 || rows->type == FPDB_TYPE_INT32, FPDB_ERR_INVALID,
             std::string(what) + ": row numbers must be INT64 or INT32");
  FPDB_CHECK(rows->length >= num_rows && rows->offset >= 0, FPDB_ERR_INVALID,
             std::string(what) + ": row number column too short");
  FPDB_CHECK(num_rows <= 0 || rows->values, FPDB_ERR_INVALID, std::string(what) + ": NULL row numbers");
  r->values = static_cast<const uint8_t*>(rows->values);
  r->validity = rows->validity;
  r->offset = rows->offset;
  r->wide = rows->type == FPDB_TYPE_INT64;
  r->limit = limit;
  return FPDB_OK;
}

// Bytes of the rows `rows` names; fails if a row number is out of range.
fpdb_status take_bytes(Context* c, const Utf8In& in, const RowsIn& rows, int64_t n, int64_t* bytes) {
  if (n == 0) {
    *bytes = 0;
    return FPDB_OK;
  }
  const int per = take_per_thread(n);
  const int blocks = (int)ceil_div(n, (int64_t)per * kStrThreads);
  int* err = ctx_err_flag(c);
  hipLaunchKernelGGL(take_len_kernel, dim3(blocks), dim3(kStrThreads), 0, c->stream, in, rows, n, per,
                     c->d_partials, err);
  FPDB_HIP(hipGetLastError());
  FPDB_TRY(launch_sum_partials(c, blocks, c->d_counter));
  return read_count(c, bytes) == FPDB_OK
             ? FPDB_OK
             : fail(FPDB_ERR_INVALID, "utf8 take: a row number is past the end of the column");
}

// Writes the rows `rows` names at output rows out_row.., their bytes from
// out_offs[out_row] = base_bytes on; room_bytes is the room past base_bytes.
fpdb_status take_into(Context* c, const Utf8In& in, const RowsIn& rows, int64_t n, int32_t* out_offs,
                      uint8_t* out_data, uint8_t* out_valid, int64_t out_row, int64_t base_bytes,
                      int64_t room_bytes, int64_t* written) {
  if (n == 0) {
    *written = 0;
    return FPDB_OK;
  }
  const int per = take_per_thread(n);
  const int blocks = (int)ceil_div(n, (int64_t)per * kStrThreads);
  int* err = ctx_err_flag(c);
  hipLaunchKernelGGL(take_len_kernel, dim3(blocks), dim3(kStrThreads), 0, c->stream, in, rows, n, per,
                     c->d_partials, err);
  FPDB_HIP(hipGetLastError());
  hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(1024), 0, c->stream, c->d_partials, blocks, c->d_counter);
  FPDB_HIP(hipGetLastError());
  int64_t total = 0;
  if (read_count(c, &total) != FPDB_OK)
    return fail(FPDB_ERR_INVALID, "utf8 take: a row number is past the end of the column");
  FPDB_CHECK(total <= room_bytes, FPDB_ERR_CAPACITY,
             "utf8 take: " + std::to_string(total) + " bytes do not fit the " + std::to_string(room_bytes) +
                 " bytes of room in the output");
  FPDB_CHECK(base_bytes + total <= INT32_MAX, FPDB_ERR_CAPACITY,
             "utf8 take: more than 2^31 - 1 bytes in one utf8 column (int32 offsets)");
  hipLaunchKernelGGL(take_offsets_kernel, dim3(blocks), dim3(kStrThreads), 0, c->stream, in, rows, n, per,
                     c->d_partials, out_offs, out_valid, out_row, base_bytes, err);
  FPDB_HIP(hipGetLastError());
  hipLaunchKernelGGL(take_copy_kernel, dim3(grid_for(c, n, kStrThreads)), dim3(kStrThreads), 0, c->stream, in,
                     rows, n, out_offs, out_data, out_row, err);
  FPDB_HIP(hipGetLastError());
  *written = total;
  return FPDB_OK;
}

// ---- comparison with a constant ------------------------------------------------------

struct ConstStr {
  uint8_t b[kMaxConstLen];
  int32_t len;
};

__global__ __launch_bounds__(256) void compare_utf8_kernel(Utf8In u, ConstStr cs, int32_t op, bool for_or,
                                                          const uint8_t* and_bm, int64_t n, uint8_t* bitmap,
                                                          unsigned long long* partials) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int64_t nchunks = (n + 63) >> 6;
  uint64_t count = 0;
  for (int64_t ch = wave; ch < nchunks; ch += nwaves) {
    const int64_t r = ch * 64 + lane;
    bool sel = false;
    if (r < n) {
      int32_t s;
      bool valid;
      const int32_t len = span_of(u, r, &s, &valid);
      if (valid) {
        const int32_t m = len < cs.len ? len : cs.len;
        int cmp = 0;
        for (int32_t i = 0; i < m && cmp == 0; ++i) {
          const uint8_t a = u.data[s + i], b = cs.b[i];
          cmp = a < b ? -1 : (a > b ? 1 : 0);
        }
        if (cmp == 0) cmp = len < cs.len ? -1 : (len > cs.len ? 1 : 0);
        switch (op) {
          case FPDB_CMP_LT: sel = cmp < 0; break;
          case FPDB_CMP_LE: sel = cmp <= 0; break;
          case FPDB_CMP_GT: sel = cmp > 0; break;
          case FPDB_CMP_GE: sel = cmp >= 0; break;
          case FPDB_CMP_EQ: sel = cmp == 0; break;
          default: sel = cmp != 0; break;
        }
      }
      if (and_bm) sel = for_or ? (sel || bitmap_bit(and_bm, r)) : (sel && bitmap_bit(and_bm, r));
    }
    const uint64_t word = __ballot(sel);
    write_bitmap_word(bitmap, ch, word, n, lane);
    count += __popcll(word);
  }
  if (partials) block_count_store(partials, lane == 0 ? count : 0);
}

// ---- key dictionary ------------------------------------------------------------------

enum KeyKind : int32_t { kKindSigned = 0, kKindUnsigned = 1, kKindFloat = 2, kKindUtf8 = 3 };

// Key columns of an input batch or of the dictionary's stored tuples.
struct KeySrc {
  const uint8_t* values[kDictMaxKeys];    // fixed-width values, or utf8 int32 offsets
  const uint8_t* data[kDictMaxKeys];      // utf8 bytes
  const uint8_t* validity[kDictMaxKeys];
  int64_t offset[kDictMaxKeys];
  int32_t width[kDictMaxKeys];            // 0 for utf8
  int32_t kind[kDictMaxKeys];
  int32_t n;
};

// A fixed-width key as a 64-bit value comparable across widths: integers
// sign- or zero-extended, floats as the bits of a double with -0.0 read as 0.0.
__device__ __forceinline__ uint64_t key_word(const KeySrc& s, int c, int64_t k) {
  if (s.kind[c] == kKindSigned) return (uint64_t)load_sext(s.values[c], s.width[c], k);
  if (s.kind[c] == kKindUnsigned) return load_zext(s.values[c], s.width[c], k);
  double d = s.width[c] == 4 ? (double)reinterpret_cast<const float*>(s.values[c])[k]
                             : reinterpret_cast<const double*>(s.values[c])[k];
  if (d == 0.0) d = 0.0;
  return (uint64_t)__double_as_longlong(d);
}

// Body step of MurmurHash3 x64 (Austin Appleby, public domain): one 64-bit
// word mixed into the running hash.
__device__ __forceinline__ uint64_t mix_word(uint64_t h, uint64_t v) {
  v *= 0x87c37b91114253d5ULL;
  v = (v << 31) | (v >> 33);
  v *= 0x4cf5ad432745937fULL;
  h ^= v;
  h = (h << 27) | (h >> 37);
  return h * 5 + 0x52dce729;
}

__device__ __forceinline__ bool any_null(const KeySrc& s, int64_t r) {
  for (int c = 0; c < s.n; ++c)
    if (!valid_bit(s.validity[c], r + s.offset[c])) return true;
  return false;
}

__device__ uint64_t tuple_hash(const KeySrc& s, int64_t r, int hash_bits) {
  uint64_t h = 0x9ae16a3b2f90404fULL;
  for (int c = 0; c < s.n; ++c) {
    const int64_t k = r + s.offset[c];
    if (!valid_bit(s.validity[c], k)) {
      h = mix_word(h, 0x5bd1e9955bd1e995ULL);
      continue;
    }
    if (s.kind[c] != kKindUtf8) {
      h = mix_word(h, key_word(s, c, k));
      continue;
    }
    const int32_t* offs = reinterpret_cast<const int32_t*>(s.values[c]);
    const int32_t b = offs[k], e = offs[k + 1];
    const uint8_t* p = s.data[c];
    int32_t i = b;
    for (; i + 8 <= e; i += 8) {
      uint64_t w = 0;
#pragma unroll
      for (int j = 0; j < 8; ++j) w |= (uint64_t)p[i + j] << (8 * j);
      h = mix_word(h, w);
    }
    uint64_t w = 0;
    for (int j = 0; i + j < e; ++j) w |= (uint64_t)p[i + j] << (8 * j);
    h = mix_word(mix_word(h, w), (uint64_t)(e - b));
  }
  h = fmix64(h);
  if (hash_bits < 64) h &= (1ULL << hash_bits) - 1;
  return h == kEmptySlot ? kEmptySlot - 1 : h;
}

__device__ bool tuple_equal(const KeySrc& a, int64_t ra, const KeySrc& b, int64_t rb) {
  for (int c = 0; c < a.n; ++c) {
    const int64_t ka = ra + a.offset[c], kb = rb + b.offset[c];
    const bool va = valid_bit(a.validity[c], ka), vb = valid_bit(b.validity[c], kb);
    if (va != vb) return false;
    if (!va) continue;
    if (a.kind[c] != kKindUtf8) {
      if (key_word(a, c, ka) != key_word(b, c, kb)) return false;
      continue;
    }
    const int32_t* oa = reinterpret_cast<const int32_t*>(a.values[c]);
    const int32_t* ob = reinterpret_cast<const int32_t*>(b.values[c]);
    const int32_t sa = oa[ka], la = oa[ka + 1] - sa, sb = ob[kb], lb = ob[kb + 1] - sb;
    if (la != lb) return false;
    for (int32_t i = 0; i < la; ++i)
      if (a.data[c][sa + i] != b.data[c][sb + i]) return false;
  }
  return true;
}

struct Table {
  unsigned long long* hash;   // kEmptySlot when free
  unsigned long long* rep;    // smallest row that claimed the slot in the current pass
  int32_t* id;                // -1 until assigned
  uint64_t mask;
};

__global__ void dict_table_init_kernel(Table t, uint64_t cap) {
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < cap; s += (uint64_t)gridDim.x * blockDim.x) {
    t.hash[s] = kEmptySlot;
    t.rep[s] = ~0ULL;
    t.id[s] = -1;
  }
}

// Re-inserts the stored tuples (distinct by construction) into a fresh table.
__global__ void dict_rehash_kernel(KeySrc store, int64_t num_ids, int hash_bits, Table t, int* err) {
  for (int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; id < num_ids;
       id += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t h = tuple_hash(store, id, hash_bits);
    uint64_t s = h & t.mask;
    bool placed = false;
    for (uint64_t p = 0; p <= t.mask; ++p, s = (s + 1) & t.mask) {
      if (atomicCAS(&t.hash[s], kEmptySlot, (unsigned long long)h) == kEmptySlot) {
        t.id[s] = (int32_t)id;
        placed = true;
        break;
      }
    }
    if (!placed) *err = 1;
  }
}

__global__ void dict_hash_kernel(KeySrc in, int64_t n, int null_is_key, int hash_bits, uint64_t* hashes) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
    hashes[r] = (!null_is_key && any_null(in, r)) ? kEmptySlot : tuple_hash(in, r, hash_bits);
}

// tent[r]: the slot row r joined in this pass (>= 0), -1 when its id is
// already final, -2 while it waits for the next pass.
__global__ void dict_insert_kernel(KeySrc in, const int64_t* list, int64_t m, const uint64_t* hashes, int insert,
                                   Table t, KeySrc store, int32_t* ids, int64_t* tent, int* err) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = list ? list[i] : i;
    const uint64_t h = hashes[r];
    int32_t id_out = -1;
    int64_t slot_out = -1;
    if (h != kEmptySlot) {
      uint64_t s = h & t.mask;
      bool done = false;
      for (uint64_t p = 0; p <= t.mask && !done; ++p, s = (s + 1) & t.mask) {
        uint64_t k = t.hash[s];
        if (k == kEmptySlot) {
          if (!insert) {
            done = true;
            break;
          }
          const uint64_t prev = atomicCAS(&t.hash[s], kEmptySlot, (unsigned long long)h);
          if (prev == kEmptySlot || prev == h) {
            // claimed now, by this row or by another row of this pass
            atomicMin(&t.rep[s], (unsigned long long)r);
            slot_out = (int64_t)s;
            done = true;
            break;
          }
          k = prev;
        }
        if (k != h) continue;
        const int32_t sid = t.id[s];
        if (sid >= 0) {
          if (tuple_equal(in, r, store, sid)) {
            id_out = sid;
            done = true;
          }
          continue;
        }
        if (!insert) continue;
        atomicMin(&t.rep[s], (unsigned long long)r);
        slot_out = (int64_t)s;
        done = true;
      }
      if (!done) *err = 1;    // no free slot: the table is sized so this cannot happen
    }
    ids[r] = id_out;
    tent[r] = slot_out;
  }
}

__global__ void dict_verify_kernel(KeySrc in, const int64_t* list, int64_t m, int64_t* tent, const Table t,
                                   uint8_t* first, int64_t* retry, unsigned long long* retry_count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = list ? list[i] : i;
    const int64_t s = tent[r];
    if (s < 0) continue;
    const int64_t rep = (int64_t)t.rep[s];
    if (rep == r) {
      first[r] = 1;
    } else if (!tuple_equal(in, r, in, rep)) {
      tent[r] = -2;
      retry[atomicAdd(retry_count, 1ULL)] = r;
    }
  }
}

// The representatives in row order take ids base_id..; their fixed-width key
// values and validity go to the stored columns (utf8 values are appended
// by the take kernels).
struct StoreOut {
  uint8_t* values[kDictMaxKeys];
  uint8_t* validity[kDictMaxKeys];
  int32_t width[kDictMaxKeys];
  int32_t n;
};

__global__ void dict_assign_kernel(KeySrc in, const int64_t* firsts, int64_t f, const int64_t* tent, Table t,
                                   uint8_t* first, int64_t base_id, StoreOut so) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < f; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = firsts[j];
    const int64_t id = base_id + j;
    t.id[tent[r]] = (int32_t)id;
    first[r] = 0;
    for (int c = 0; c < so.n; ++c) {
      const int64_t k = r + in.offset[c];
      const bool v = valid_bit(in.validity[c], k);
      if (v) atomicOr(reinterpret_cast<unsigned int*>(so.validity[c]) + (id >> 5), 1u << (id & 31));
      if (so.width[c] > 0) store_w(so.values[c], so.width[c], id, v ? load_zext(in.values[c], in.width[c], k) : 0);
    }
  }
}

__global__ void dict_finish_kernel(const int64_t* list, int64_t m, const int64_t* tent, const Table t, int32_t* ids) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = list ? list[i] : i;
    const int64_t s = tent[r];
    if (s >= 0) ids[r] = t.id[s];
  }
}

// ids -> the caller's INT32 column (and its validity: bit set where id >= 0)
__global__ void dict_ids_out_kernel(const int32_t* ids, int64_t n, int32_t* out, uint8_t* out_valid) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t ch = wave; ch < (n + 63) >> 6; ch += nwaves) {
    const int64_t r = ch * 64 + lane;
    const int32_t id = r < n ? ids[r] : -1;
    if (r < n) out[r] = id;
    if (out_valid) write_bitmap_word(out_valid, ch, __ballot(r < n && id >= 0), n, lane);
  }
}

__global__ void decode_fixed_kernel(const uint8_t* values, const uint8_t* validity, int32_t width, RowsIn ids,
                                    int64_t n, uint8_t* out, uint8_t* out_valid, int64_t out_offset, int* err) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t id = rows_at(ids, i, err);
    const bool v = id >= 0 && valid_bit(validity, id);
    store_w(out, width, out_offset + i, v ? load_zext(values, width, id) : 0);
    if (out_valid && v) {
      const int64_t b = out_offset + i;
      atomicOr(reinterpret_cast<unsigned int*>(out_valid) + (b >> 5), 1u << (b & 31));
    }
  }
}

__global__ void iota_kernel(int64_t* out, int64_t n, int64_t start) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = start + i;
}

}  // namespace
}  // namespace fpdb

using namespace fpdb;

struct DictColumn {
  int32_t type = 0, kind = 0, width = 0;
  uint8_t* values = nullptr;     // fixed values (id_cap * width) or utf8 offsets (id_cap + 1)
  uint8_t* validity = nullptr;   // id_cap bits, 4-byte words
  uint8_t* data = nullptr;       // utf8 bytes
  int64_t data_cap = 0, data_used = 0;
};

struct fpdb_keydict {
  Context* c = nullptr;
  int n = 0;
  int hash_bits = 64;
  DictColumn col[kDictMaxKeys];
  int64_t num_ids = 0, id_cap = 0;
  uint64_t cap = 0;              // table slots (power of two), 0 before the first insert
  Table t{};
};

namespace {

KeySrc store_src(const fpdb_keydict* d) {
  KeySrc s;
  std::memset(&s, 0, sizeof(s));
  s.n = d->n;
  for (int i = 0; i < d->n; ++i) {
    s.values[i] = d->col[i].values;
    s.data[i] = d->col[i].data;
    s.validity[i] = d->col[i].validity;
    s.width[i] = d->col[i].width;
    s.kind[i] = d->col[i].kind;
  }
  return s;
}

int32_t kind_of(int32_t type) {
  if (type == FPDB_TYPE_UTF8) return kKindUtf8;
  if (type_is_float(type)) return kKindFloat;
  return type_is_signed_int(type) ? kKindSigned : kKindUnsigned;
}

fpdb_status key_src(const fpdb_keydict* d, const fpdb_column* keys, int num_keys, int64_t n, bool exact_types,
                    KeySrc* s) {
  FPDB_CHECK(num_keys == d->n, FPDB_ERR_INVALID,
             "key dictionary: " + std::to_string(num_keys) + " key columns, the dictionary has " +
                 std::to_string(d->n));
  FPDB_CHECK(keys || n == 0, FPDB_ERR_INVALID, "key dictionary: NULL key columns");
  std::memset(s, 0, sizeof(*s));
  s->n = num_keys;
  for (int i = 0; i < num_keys; ++i) {
    const fpdb_column& k = keys[i];
    const int32_t kind = kind_of(k.type);
    FPDB_CHECK(k.type == FPDB_TYPE_UTF8 || type_width(k.type) > 0, FPDB_ERR_INVALID,
               "key dictionary: unsupported key column type");
    FPDB_CHECK(kind == d->col[i].kind, FPDB_ERR_INVALID,
               "key dictionary: key column " + std::to_string(i) + " differs in kind from the dictionary's " +
                   "(signed, unsigned, float and utf8 do not mix)");
    FPDB_CHECK(!exact_types || k.type == d->col[i].type, FPDB_ERR_INVALID,
               "key dictionary: inserted key column " + std::to_string(i) + " must have the dictionary's type");
    FPDB_CHECK(k.length >= n && k.offset >= 0, FPDB_ERR_INVALID, "key dictionary: key column too short");
    FPDB_CHECK(n == 0 || k.values, FPDB_ERR_INVALID, "key dictionary: NULL key values");
    FPDB_CHECK(n == 0 || kind != kKindUtf8 || k.data || k.data_bytes == 0, FPDB_ERR_INVALID,
               "key dictionary: NULL utf8 data");
    s->values[i] = static_cast<const uint8_t*>(k.values);
    s->data[i] = static_cast<const uint8_t*>(k.data);
    s->validity[i] = k.validity;
    s->offset[i] = k.offset;
    s->width[i] = kind == kKindUtf8 ? 0 : type_width(k.type);
    s->kind[i] = kind;
  }
  return FPDB_OK;
}

uint64_t pow2_at_least(uint64_t x) {
  uint64_t p = 1024;
  while (p < x) p <<= 1;
  return p;
}

fpdb_status table_alloc(fpdb_keydict* d, uint64_t cap) {
  Context* c = d->c;
  Table t{};
  void* p[3] = {nullptr, nullptr, nullptr};
  FPDB_TRY(ctx_alloc(c, cap * 8, &p[0]));
  if (ctx_alloc(c, cap * 8, &p[1]) != FPDB_OK || ctx_alloc(c, cap * 4, &p[2]) != FPDB_OK) {
    for (void* q : p)
      if (q) ctx_free(c, q);
    return fail(FPDB_ERR_BUDGET, "key dictionary: table of " + std::to_string(cap) + " slots does not fit");
  }
  t.hash = static_cast<unsigned long long*>(p[0]);
  t.rep = static_cast<unsigned long long*>(p[1]);
  t.id = static_cast<int32_t*>(p[2]);
  t.mask = cap - 1;
  hipLaunchKernelGGL(dict_table_init_kernel, dim3(grid_for(c, (int64_t)cap, 256)), dim3(256), 0, c->stream, t, cap);
  FPDB_HIP(hipGetLastError());
  if (d->cap) {
    ctx_free(c, d->t.hash);
    ctx_free(c, d->t.rep);
    ctx_free(c, d->t.id);
  }
  d->t = t;
  d->cap = cap;
  if (d->num_ids) {
    hipLaunchKernelGGL(dict_rehash_kernel, dim3(grid_for(c, d->num_ids, 256)), dim3(256), 0, c->stream,
                       store_src(d), d->num_ids, d->hash_bits, d->t, ctx_err_flag(c));
    FPDB_HIP(hipGetLastError());
  }
  return FPDB_OK;
}

// Grows a device buffer to `bytes`, keeping its first `keep` bytes and zeroing
// the rest.
fpdb_status grow_buffer(Context* c, uint8_t** buf, uint64_t keep, uint64_t bytes) {
  void* p = nullptr;
  FPDB_TRY(ctx_alloc(c, bytes, &p));
  FPDB_TRY(launch_fill(c, p, 0, bytes));
  if (*buf && keep) FPDB_HIP(hipMemcpyAsync(p, *buf, keep, hipMemcpyDeviceToDevice, c->stream));
  if (*buf) ctx_free(c, *buf);
  *buf = static_cast<uint8_t*>(p);
  return FPDB_OK;
}

fpdb_status store_reserve_ids(fpdb_keydict* d, int64_t ids) {
  if (ids <= d->id_cap) return FPDB_OK;
  const int64_t cap = std::max<int64_t>(ids, std::max<int64_t>(1024, d->id_cap * 2));
  Context* c = d->c;
  for (int i = 0; i < d->n; ++i) {
    DictColumn& k = d->col[i];
    const uint64_t vbytes_old = (uint64_t)ceil_div(d->id_cap, 32) * 4, vbytes = (uint64_t)ceil_div(cap, 32) * 4;
    FPDB_TRY(grow_buffer(c, &k.validity, d->id_cap ? vbytes_old : 0, vbytes));
    if (k.kind == kKindUtf8) {
      FPDB_TRY(grow_buffer(c, &k.values, d->id_cap ? (uint64_t)(d->id_cap + 1) * 4 : 0, (uint64_t)(cap + 1) * 4));
    } else {
      FPDB_TRY(grow_buffer(c, &k.values, (uint64_t)d->id_cap * k.width, (uint64_t)cap * k.width));
    }
  }
  d->id_cap = cap;
  return FPDB_OK;
}

void free_all(fpdb_keydict* d) {
  Context* c = d->c;
  for (int i = 0; i < d->n; ++i) {
    if (d->col[i].values) ctx_free(c, d->col[i].values);
    if (d->col[i].validity) ctx_free(c, d->col[i].validity);
    if (d->col[i].data) ctx_free(c, d->col[i].data);
  }
  if (d->cap) {
    ctx_free(c, d->t.hash);
    ctx_free(c, d->t.rep);
    ctx_free(c, d->t.id);
  }
}

// Per-call device scratch, freed on every return path.
struct Scratch {
  Context* c;
  std::vector<void*> bufs;
  explicit Scratch(Context* ctx) : c(ctx) {}
  ~Scratch() {
    for (void* p : bufs) ctx_free(c, p);
  }
  template <class T>
  fpdb_status get(uint64_t count, T** out) {
    void* p = nullptr;
    FPDB_TRY(ctx_alloc(c, std::max<uint64_t>(count, 1) * sizeof(T), &p));
    bufs.push_back(p);
    *out = static_cast<T*>(p);
    return FPDB_OK;
  }
};

}  // namespace

extern "C" {

fpdb_status fpdb_utf8_take_bytes(fpdb_context* ctx, const fpdb_column* in, const fpdb_column* rows,
                                 int64_t num_rows, int64_t* bytes) {
  FPDB_CHECK(ctx && bytes, FPDB_ERR_INVALID, "NULL argument");
  FPDB_CHECK(num_rows >= 0, FPDB_ERR_INVALID, "utf8 take: negative row count");
  Utf8In u;
  RowsIn r;
  FPDB_TRY(utf8_view(in, num_rows, &u, "utf8 take"));
  FPDB_TRY(rows_view(rows, num_rows, in->length, &r, "utf8 take"));
  return take_bytes(as_ctx(ctx), u, r, num_rows, bytes);
}

fpdb_status fpdb_utf8_take(fpdb_context* ctx, const fpdb_column* in, const fpdb_column* rows, int64_t num_rows,
                           fpdb_column* out) {
  FPDB_CHECK(ctx && out, FPDB_ERR_INVALID, "NULL argument");
  FPDB_CHECK(num_rows >= 0, FPDB_ERR_INVALID, "utf8 take: negative row count");
  Utf8In u;
  RowsIn r;
  FPDB_TRY(utf8_view(in, num_rows, &u, "utf8 take"));
  FPDB_TRY(rows_view(rows, num_rows, in->length, &r, "utf8 take"));
  FPDB_CHECK(out->type == FPDB_TYPE_UTF8 && out->offset == 0 && out->values, FPDB_ERR_INVALID,
             "utf8 take: the output must be a utf8 column at offset 0 with an offsets buffer");
  FPDB_CHECK(out->length >= num_rows, FPDB_ERR_INVALID, "utf8 take: output column too short");
  FPDB_CHECK(out->data || out->data_bytes == 0, FPDB_ERR_INVALID, "utf8 take: NULL output data");
  FPDB_CHECK(out->validity || (!in->validity && !rows->validity), FPDB_ERR_INVALID,
             "utf8 take: a nullable input or row column needs an output validity bitmap");
  Context* c = as_ctx(ctx);
  if (out->validity) FPDB_TRY(launch_fill(c, out->validity, 0, (uint64_t)ceil_div(num_rows, 32) * 4));
  if (num_rows == 0) return launch_fill(c, out->values, 0, 4);
  int64_t written = 0;
  return take_into(c, u, r, num_rows, static_cast<int32_t*>(out->values), static_cast<uint8_t*>(out->data),
                   out->validity, 0, 0, out->data_bytes, &written);
}

fpdb_status fpdb_iota_int64(fpdb_context* ctx, int64_t* d_out, int64_t n, int64_t start) {
  FPDB_CHECK(ctx && (d_out || n == 0), FPDB_ERR_INVALID, "NULL argument");
  FPDB_CHECK(n >= 0, FPDB_ERR_INVALID, "iota: negative count");
  if (n == 0) return FPDB_OK;
  Context* c = as_ctx(ctx);
  hipLaunchKernelGGL(iota_kernel, dim3(grid_for(c, n, 256)), dim3(256), 0, c->stream, d_out, n, start);
  FPDB_HIP(hipGetLastError());
  return FPDB_OK;
}

fpdb_status fpdb_compare_bitmap_utf8(fpdb_context* ctx, const fpdb_column* column, int op, const char* constant,
                                     int64_t constant_len, const uint8_t* d_and_bitmap, int64_t num_rows,
                                     uint8_t* d_bitmap, int64_t* d_num_selected) {
  FPDB_CHECK(ctx && column && d_bitmap && (constant || constant_len == 0), FPDB_ERR_INVALID, "NULL argument");
  FPDB_CHECK(constant_len >= 0 && constant_len <= kMaxConstLen, FPDB_ERR_INVALID,
             "utf8 compare: the constant may hold at most " + std::to_string(kMaxConstLen) + " bytes");
  const bool for_or = (op & FPDB_CMP_OR) != 0;
  op &= ~FPDB_CMP_OR;
  FPDB_CHECK(op >= FPDB_CMP_LT && op <= FPDB_CMP_NE, FPDB_ERR_INVALID, "unknown comparison");
  FPDB_CHECK(!for_or || d_and_bitmap, FPDB_ERR_INVALID, "FPDB_CMP_OR needs a bitmap to combine with");
  FPDB_CHECK(column->length >= num_rows, FPDB_ERR_INVALID, "utf8 compare: column too short");
  Utf8In u;
  FPDB_TRY(utf8_view(column, num_rows, &u, "utf8 compare"));
  Context* c = as_ctx(ctx);
  if (num_rows == 0) return d_num_selected ? launch_fill(c, d_num_selected, 0, 8) : FPDB_OK;
  ConstStr cs;
  std::memset(&cs, 0, sizeof(cs));
  if (constant_len) std::memcpy(cs.b, constant, (size_t)constant_len);
  cs.len = (int32_t)constant_len;
  ProfScope ps(c, "compare_utf8");
  const int grid = grid_for(c, num_rows, 256, 16);
  hipLaunchKernelGGL(compare_utf8_kernel, dim3(grid), dim3(256), 0, c->stream, u, cs, op, for_or, d_and_bitmap,
                     num_rows, d_bitmap, d_num_selected ? c->d_partials : nullptr);
  FPDB_HIP(hipGetLastError());
  if (d_num_selected) FPDB_TRY(launch_sum_partials(c, grid, d_num_selected));
  return FPDB_OK;
}

fpdb_status fpdb_keydict_create(fpdb_context* ctx, const int32_t* key_types, int num_keys, int hash_bits,
                                fpdb_keydict** out) {
  FPDB_CHECK(ctx && key_types && out, FPDB_ERR_INVALID, "NULL argument");
  FPDB_CHECK(num_keys >= 1 && num_keys <= kDictMaxKeys, FPDB_ERR_INVALID,
             "key dictionary: 1 to " + std::to_string(kDictMaxKeys) + " key columns");
  FPDB_CHECK(hash_bits >= 1 && hash_bits <= 64, FPDB_ERR_INVALID, "key dictionary: hash_bits must be in [1, 64]");
  for (int i = 0; i < num_keys; ++i)
    FPDB_CHECK(key_types[i] == FPDB_TYPE_UTF8 || type_width(key_types[i]) > 0, FPDB_ERR_INVALID,
               "key dictionary: unsupported key type");
  auto* d = new fpdb_keydict();
  d->c = as_ctx(ctx);
  d->n = num_keys;
  d->hash_bits = hash_bits;
  for (int i = 0; i < num_keys; ++i) {
    d->col[i].type = key_types[i];
    d->col[i].kind = kind_of(key_types[i]);
    d->col[i].width = key_types[i] == FPDB_TYPE_UTF8 ? 0 : type_width(key_types[i]);
  }
  *out = d;
  return FPDB_OK;
}

fpdb_status fpdb_keydict_destroy(fpdb_keydict* d) {
  if (!d) return FPDB_OK;
  free_all(d);
  delete d;
  return FPDB_OK;
}

fpdb_status fpdb_keydict_size(fpdb_keydict* d, int64_t* num_ids) {
  FPDB_CHECK(d && num_ids, FPDB_ERR_INVALID, "NULL argument");
  *num_ids = d->num_ids;
  return FPDB_OK;
}

fpdb_status fpdb_keydict_encode(fpdb_context* ctx, fpdb_keydict* d, const fpdb_column* keys, int num_keys,
                                int64_t n, int insert, int null_is_key, fpdb_column* ids, int64_t* num_ids) {
  FPDB_CHECK(ctx && d && ids, FPDB_ERR_INVALID, "NULL argument");
  FPDB_CHECK(as_ctx(ctx) == d->c, FPDB_ERR_INVALID, "key dictionary: used with another context");
  FPDB_CHECK(n >= 0, FPDB_ERR_INVALID, "key dictionary: negative row count");
  FPDB_CHECK(ids->type == FPDB_TYPE_INT32 && ids->offset == 0 && ids->length >= n && (ids->values || n == 0),
             FPDB_ERR_INVALID, "key dictionary: ids must be an INT32 column at offset 0 with room for every row");
  Context* c = d->c;
  KeySrc in;
  FPDB_TRY(key_src(d, keys, num_keys, n, insert != 0, &in));
  if (n > 0) {
    Scratch sc(c);
    uint64_t* hashes;
    int64_t* tent;
    int32_t* idbuf;
    FPDB_TRY(sc.get(n, &hashes));
    FPDB_TRY(sc.get(n, &tent));
    FPDB_TRY(sc.get(n, &idbuf));
    const int grid = grid_for(c, n, 256);
    hipLaunchKernelGGL(dict_hash_kernel, dim3(grid), dim3(256), 0, c->stream, in, n, null_is_key, d->hash_bits,
                       hashes);
    FPDB_HIP(hipGetLastError());
    if (insert) {
      FPDB_CHECK(d->num_ids + n < INT32_MAX, FPDB_ERR_CAPACITY, "key dictionary: more than 2^31 - 1 ids");
      const uint64_t want = pow2_at_least(2 * (uint64_t)(d->num_ids + n));
      if (d->cap < want) FPDB_TRY(table_alloc(d, want));
    }
    if (d->cap == 0) {
      // nothing inserted yet: every lookup misses
      FPDB_TRY(launch_fill(c, idbuf, 0xFF, (uint64_t)n * 4));
    } else if (!insert) {
      hipLaunchKernelGGL(dict_insert_kernel, dim3(grid), dim3(256), 0, c->stream, in, (const int64_t*)nullptr, n,
                         hashes, 0, d->t, store_src(d), idbuf, tent, ctx_err_flag(c));
      FPDB_HIP(hipGetLastError());
    } else {
      uint8_t* first;
      int64_t *list_a, *list_b, *firsts;
      FPDB_TRY(sc.get(n, &first));
      FPDB_TRY(sc.get(n, &list_a));
      FPDB_TRY(sc.get(n, &list_b));
      FPDB_TRY(sc.get(n, &firsts));
      FPDB_TRY(launch_fill(c, first, 0, (uint64_t)n));
      const int64_t* list = nullptr;
      int64_t m = n;
      for (int64_t pass = 0;; ++pass) {
        FPDB_CHECK(pass <= n, FPDB_ERR_INTERNAL, "key dictionary: insertion did not settle");
        int64_t* retry = (pass & 1) ? list_a : list_b;
        const int pg = grid_for(c, m, 256);
        hipLaunchKernelGGL(dict_insert_kernel, dim3(pg), dim3(256), 0, c->stream, in, list, m, hashes, 1, d->t,
                           store_src(d), idbuf, tent, ctx_err_flag(c));
        FPDB_HIP(hipGetLastError());
        FPDB_TRY(launch_fill(c, c->d_counter + 3, 0, 8));
        hipLaunchKernelGGL(dict_verify_kernel, dim3(pg), dim3(256), 0, c->stream, in, list, m, tent, d->t, first,
                           retry, reinterpret_cast<unsigned long long*>(c->d_counter + 3));
        FPDB_HIP(hipGetLastError());
        // the representatives, in row order
        ColSet none;
        OutColSet onone;
        std::memset(&none, 0, sizeof(none));
        std::memset(&onone, 0, sizeof(onone));
        onone.capacity = INT64_MAX;
        int64_t f = 0;
        FPDB_TRY((run_compact<PredByteFlag, false>(c, PredByteFlag{first, 1}, none, onone, n, nullptr, firsts, &f,
                                                   nullptr, "keydict_firsts")));
        if (f > 0) {
          const int64_t base = d->num_ids;
          FPDB_TRY(store_reserve_ids(d, base + f));
          StoreOut so;
          std::memset(&so, 0, sizeof(so));
          so.n = d->n;
          for (int i = 0; i < d->n; ++i) {
            so.values[i] = d->col[i].values;
            so.validity[i] = d->col[i].validity;
            so.width[i] = d->col[i].width;
          }
          hipLaunchKernelGGL(dict_assign_kernel, dim3(grid_for(c, f, 256)), dim3(256), 0, c->stream, in, firsts, f,
                             tent, d->t, first, base, so);
          FPDB_HIP(hipGetLastError());
          for (int i = 0; i < d->n; ++i) {
            DictColumn& k = d->col[i];
            if (k.kind != kKindUtf8) continue;
            Utf8In u{reinterpret_cast<const int32_t*>(in.values[i]), in.data[i], in.validity[i], in.offset[i]};
            RowsIn rr{reinterpret_cast<const uint8_t*>(firsts), nullptr, 0, 1, INT64_MAX};
            int64_t bytes = 0;
            FPDB_TRY(take_bytes(c, u, rr, f, &bytes));
            const int64_t need = k.data_used + bytes;
            FPDB_CHECK(need <= INT32_MAX, FPDB_ERR_CAPACITY,
                       "key dictionary: more than 2^31 - 1 bytes of strings in one key column");
            if (need > k.data_cap) {
              const int64_t cap =
                  std::min<int64_t>(INT32_MAX, std::max<int64_t>(need, std::max<int64_t>(4096, 2 * k.data_cap)));
              FPDB_TRY(grow_buffer(c, &k.data, (uint64_t)k.data_used, (uint64_t)cap));
              k.data_cap = cap;
            }
            // validity bits were set by dict_assign_kernel; the take writes offsets and bytes only
            int64_t written = 0;
            FPDB_TRY(take_into(c, u, rr, f, reinterpret_cast<int32_t*>(k.values), k.data, nullptr, base,
                               k.data_used, k.data_cap - k.data_used, &written));
            k.data_used += written;
          }
          d->num_ids = base + f;
        }
        hipLaunchKernelGGL(dict_finish_kernel, dim3(pg), dim3(256), 0, c->stream, list, m, tent, d->t, idbuf);
        FPDB_HIP(hipGetLastError());
        uint64_t left = 0;
        FPDB_TRY(mail_read(c, c->d_counter + 3, 1, &left));
        if (left == 0) break;
        list = retry;
        m = (int64_t)left;
      }
    }
    hipLaunchKernelGGL(dict_ids_out_kernel, dim3(grid_for(c, n, 256)), dim3(256), 0, c->stream, idbuf, n,
                       static_cast<int32_t*>(ids->values), ids->validity);
    FPDB_HIP(hipGetLastError());
    int64_t dummy = 0;
    if (read_count(c, &dummy) != FPDB_OK)
      return fail(FPDB_ERR_INTERNAL, "key dictionary: the table overflowed");
  }
  if (num_ids) *num_ids = d->num_ids;
  return FPDB_OK;
}

fpdb_status fpdb_keydict_decode_bytes(fpdb_context* ctx, fpdb_keydict* d, int key, const fpdb_column* ids,
                                      int64_t num_rows, int64_t* bytes) {
  FPDB_CHECK(ctx && d && bytes, FPDB_ERR_INVALID, "NULL argument");
  FPDB_CHECK(key >= 0 && key < d->n, FPDB_ERR_INVALID, "key dictionary: no such key column");
  FPDB_CHECK(d->col[key].kind == kKindUtf8, FPDB_ERR_INVALID, "key dictionary: key column is not utf8");
  RowsIn r;
  FPDB_TRY(rows_view(ids, num_rows, d->num_ids, &r, "key dictionary decode"));
  Utf8In u{reinterpret_cast<const int32_t*>(d->col[key].values), d->col[key].data, d->col[key].validity, 0};
  return take_bytes(as_ctx(ctx), u, r, num_rows, bytes);
}

fpdb_status fpdb_keydict_decode(fpdb_context* ctx, fpdb_keydict* d, int key, const fpdb_column* ids,
                                int64_t num_rows, fpdb_column* out) {
  FPDB_CHECK(ctx && d && out, FPDB_ERR_INVALID, "NULL argument");
  FPDB_CHECK(key >= 0 && key < d->n, FPDB_ERR_INVALID, "key dictionary: no such key column");
  FPDB_CHECK(num_rows >= 0, FPDB_ERR_INVALID, "key dictionary: negative row count");
  const DictColumn& k = d->col[key];
  FPDB_CHECK(out->type == k.type, FPDB_ERR_INVALID, "key dictionary decode: output type differs from the key's");
  FPDB_CHECK(out->length >= num_rows, FPDB_ERR_INVALID, "key dictionary decode: output column too short");
  Context* c = as_ctx(ctx);
  RowsIn r;
  FPDB_TRY(rows_view(ids, num_rows, d->num_ids, &r, "key dictionary decode"));
  if (k.kind == kKindUtf8) {
    fpdb_column in{};
    in.type = FPDB_TYPE_UTF8;
    in.length = d->num_ids;
    in.values = k.values;
    in.validity = k.validity;
    in.data = k.data;
    in.data_bytes = k.data_used;
    if (num_rows == 0 || d->num_ids == 0) {
      // no stored tuples: every row is null
      FPDB_CHECK(num_rows == 0 || out->validity, FPDB_ERR_INVALID,
                 "key dictionary decode: null rows need an output validity bitmap");
      if (out->validity) FPDB_TRY(launch_fill(c, out->validity, 0, (uint64_t)ceil_div(num_rows, 32) * 4));
      FPDB_TRY(launch_fill(c, out->values, 0, (uint64_t)(num_rows + 1) * 4));
      return FPDB_OK;
    }
    return fpdb_utf8_take(ctx, &in, ids, num_rows, out);
  }
  FPDB_CHECK((out->offset & 31) == 0, FPDB_ERR_INVALID, "key dictionary decode: output offset must be a multiple of 32");
  FPDB_CHECK(out->values || num_rows == 0, FPDB_ERR_INVALID, "key dictionary decode: NULL output values");
  if (out->validity)
    FPDB_TRY(launch_fill(c, out->validity + out->offset / 8, 0, (uint64_t)ceil_div(num_rows, 32) * 4));
  if (num_rows == 0) return FPDB_OK;
  hipLaunchKernelGGL(decode_fixed_kernel, dim3(grid_for(c, num_rows, 256)), dim3(256), 0, c->stream, k.values,
                     k.validity, k.width, r, num_rows, static_cast<uint8_t*>(out->values), out->validity,
                     out->offset, ctx_err_flag(c));
  FPDB_HIP(hipGetLastError());
  int64_t dummy = 0;
  if (read_count(c, &dummy) != FPDB_OK)
    return fail(FPDB_ERR_INVALID, "key dictionary decode: an id is past the dictionary's size");
  return FPDB_OK;
}

}  // extern "C"

// ---- utf8 columns inside the row-moving entry points ---------------------------------

namespace fpdb {
namespace {

// Appends rows [0, n) of `in` after output row out_row: offsets continue from
// out.values[out_row] (0 when out_row is 0), bytes are one contiguous copy.
// A workgroup whose bytes would pass the output's room raises the error flag
// and writes nothing.
__global__ __launch_bounds__(256) void utf8_copy_rows_kernel(Utf8In in, int64_t n, int32_t* out_offs,
                                                             uint8_t* out_data, uint8_t* out_valid, int64_t out_row,
                                                             int64_t room, int* err) {
  const int64_t base = out_row == 0 ? 0 : out_offs[out_row];
  const int32_t in0 = in.offs[in.offset], in1 = in.offs[in.offset + n];
  if (base + (int64_t)(in1 - in0) > room || base + (int64_t)(in1 - in0) > INT32_MAX) {
    if (threadIdx.x == 0) *err = 1;
    return;
  }
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t0 == 0 && out_row == 0) out_offs[0] = 0;
  for (int64_t i = t0; i < n; i += stride) {
    out_offs[out_row + i + 1] = (int32_t)(base + in.offs[in.offset + i + 1] - in0);
    if (out_valid) {
      const int64_t b = out_row + i;
      unsigned int* w = reinterpret_cast<unsigned int*>(out_valid) + (b >> 5);
      if (valid_bit(in.validity, in.offset + i))
        atomicOr(w, 1u << (b & 31));
      else
        atomicAnd(w, ~(1u << (b & 31)));
    }
  }
  for (int64_t b = t0; b < (int64_t)(in1 - in0); b += stride) out_data[base + b] = in.data[in0 + b];
}

}  // namespace

fpdb_status utf8_gather(Context* c, const fpdb_column& in, const int64_t* d_rows, int64_t n, const fpdb_column& out) {
  Utf8In u;
  FPDB_TRY(utf8_view(&in, n, &u, "gather"));
  FPDB_CHECK(out.type == FPDB_TYPE_UTF8 && out.offset == 0 && (out.values || n == 0), FPDB_ERR_INVALID,
             "gather: a utf8 output must be a utf8 column at offset 0 with an offsets buffer");
  FPDB_CHECK(out.length >= n, FPDB_ERR_INVALID, "gather: output column too short");
  RowsIn r{reinterpret_cast<const uint8_t*>(d_rows), nullptr, 0, 1, in.length};
  if (out.validity) FPDB_TRY(launch_fill(c, out.validity, 0, (uint64_t)ceil_div(n, 32) * 4));
  if (n == 0) return out.values ? launch_fill(c, out.values, 0, 4) : FPDB_OK;
  int64_t written = 0;
  return take_into(c, u, r, n, static_cast<int32_t*>(out.values), static_cast<uint8_t*>(out.data), out.validity, 0,
                   0, out.data_bytes, &written);
}

fpdb_status utf8_copy_rows(Context* c, const fpdb_column& in, int64_t n, const fpdb_column& out, int64_t out_row) {
  Utf8In u;
  FPDB_TRY(utf8_view(&in, n, &u, "copy rows"));
  FPDB_CHECK(out.type == FPDB_TYPE_UTF8 && out.offset == 0 && out.values, FPDB_ERR_INVALID,
             "copy rows: a utf8 output must be a utf8 column at offset 0 with an offsets buffer");
  FPDB_CHECK(out.validity || !in.validity, FPDB_ERR_INVALID,
             "copy rows: nullable input needs an output validity bitmap");
  if (n == 0) return out_row == 0 ? launch_fill(c, out.values, 0, 4) : FPDB_OK;
  hipLaunchKernelGGL(utf8_copy_rows_kernel, dim3(grid_for(c, n, 256)), dim3(256), 0, c->stream, u, n,
                     static_cast<int32_t*>(out.values), static_cast<uint8_t*>(out.data), out.validity, out_row,
                     out.data_bytes, ctx_err_flag(c));
  FPDB_HIP(hipGetLastError());
  int64_t dummy = 0;
  if (read_count(c, &dummy) != FPDB_OK)
    return fail(FPDB_ERR_CAPACITY, "copy rows: the utf8 bytes do not fit the output's data buffer");
  return FPDB_OK;
}

}  // namespace fpdb
