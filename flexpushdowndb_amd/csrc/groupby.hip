// Hash group-by aggregation on gfx950, with no global atomics on the row path.
//
// Every global 64-bit atomic on this part is executed memory-side at a fixed
// rate of about 24 G/s for the whole chip, whatever the table size or scope,
// so a row-at-a-time global hash table cannot stream 60M rows at HBM speed.
// Instead:
//
//   local    (one workgroup per contiguous row range, 1024 threads): a hash
//            table in LDS aggregates every row whose key finds or claims a
//            slot within a few probes (this absorbs the heavy head of skewed
//            keys).  The rest ("tail" rows) are only counted per bucket, a
//            bucket being the top bits of the key hash.  At the end the table
//            is written out slot by slot as partial records.
//   scan     bucket-major exclusive scan of the per-workgroup counts.
//   scatter  (same workgroup / row range): the LDS table is rebuilt at the same
//            slots from the partial records, so a row is a tail row exactly
//            when the local pass found no slot for it; tail rows and partial
//            records are written as records to their bucket's range.
//   compact  (one workgroup per bucket): the bucket's records, old and new,
//            are aggregated in an LDS table and written back as one record per
//            group.
//
// A record is (key, first row, accumulators...), all 8-byte words; raw rows
// and partial aggregates have the same form, so every pass combines them with
// the same operations.  Integer aggregates are exact (wrapping int64
// arithmetic is associative); FP64 sums depend on the order in which records
// are combined.
//
// Finalize: groups are emitted in order of first appearance (the order in
// which a serial grouper numbers them), ranked by a bitmap over the first-row
// positions and a prefix popcount.
#include "fpdb_internal.hpp"
#include "compact.hpp"

#include <algorithm>
#include <climits>
#include <vector>

namespace fpdb {

constexpr uint64_t kEmptyKey = ~0ULL;
constexpr int kMaxAggs = 16;          // visible + hidden accumulators
constexpr int kGroupThreads = 1024;
constexpr int kMinBuckets = 256;     // at least one compact workgroup per CU
constexpr int kMaxBuckets = 1 << 12;
constexpr int64_t kMaxBatchRows = int64_t(1) << 30;   // consume splits larger batches (32-bit batch rows in LDS)

__device__ __forceinline__ bool type_is_signed_int_d(int32_t t) {
  return t == FPDB_TYPE_INT8 || t == FPDB_TYPE_INT16 || t == FPDB_TYPE_INT32 || t == FPDB_TYPE_INT64 ||
         t == FPDB_TYPE_DATE32 || t == FPDB_TYPE_DATE64;
}

// accumulator operations
enum AccOp : int32_t {
  kAccCountAll = 0,    // += 1
  kAccCountValid = 1,  // += valid
  kAccSumI = 2,        // += int (wrapping)
  kAccSumF = 3,        // += double
  kAccMinI = 4, kAccMaxI = 5, kAccMinU = 6, kAccMaxU = 7, kAccMinF = 8, kAccMaxF = 9
};

struct AccSpec {
  int32_t op[kMaxAggs];
  int32_t input[kMaxAggs];   // index into the value column set, -1 for COUNT_ALL
  int32_t n;
};

// The group key of a row is one 64-bit word: key column j contributes
// (value - bias[j]) in bits[j] bits at bit position shift[j].  One column is
// the value itself (bias 0, 64 bits); two columns of at most 4 bytes are the
// low 32 bits of each (bias 0, 32 bits); packed keys take caller-given ranges
// and flag values outside them (check).  No key column: every row is group 0.
constexpr int kMaxGroupKeys = 4;

struct GKeySpec {
  const uint8_t* v[kMaxGroupKeys];
  int64_t off[kMaxGroupKeys];
  int64_t bias[kMaxGroupKeys];
  int32_t w[kMaxGroupKeys];
  int32_t sgn[kMaxGroupKeys];
  int32_t bits[kMaxGroupKeys];
  int32_t shift[kMaxGroupKeys];
  int32_t n;
  int32_t check;
};

__host__ __device__ __forceinline__ uint64_t key_field_mask(int bits) {
  return bits >= 64 ? ~0ULL : ((1ULL << bits) - 1);
}

__device__ __forceinline__ uint64_t group_key(const GKeySpec& k, int64_t r, bool* out_of_range) {
  uint64_t key = 0;
#pragma unroll
  for (int j = 0; j < kMaxGroupKeys; ++j) {
    if (j >= k.n) break;
    int64_t i = r + k.off[j];
    int64_t v = k.sgn[j] ? load_sext(k.v[j], k.w[j], i) : (int64_t)load_zext(k.v[j], k.w[j], i);
    uint64_t d = (uint64_t)v - (uint64_t)k.bias[j];
    uint64_t m = key_field_mask(k.bits[j]);
    if (k.check && d > m) *out_of_range = true;
    key |= (d & m) << k.shift[j];
  }
  return key;
}

__host__ __device__ __forceinline__ uint64_t acc_identity(int32_t op) {
  switch (op) {
    case kAccMinI: return (uint64_t)LLONG_MAX;
    case kAccMaxI: return (uint64_t)LLONG_MIN;
    case kAccMinU: return ~0ULL;
    case kAccMaxU: return 0ULL;
    case kAccMinF: return 0x7FF0000000000000ULL;   // +inf
    case kAccMaxF: return 0xFFF0000000000000ULL;   // -inf
    default: return 0ULL;
  }
}

// Value of row r for an accumulator, as the 8-byte operand of the atomic;
// `valid` reports the input validity.
__device__ __forceinline__ uint64_t acc_operand(const ColSet& vals, int32_t op, int32_t in, int64_t r,
                                                bool* valid) {
  if (op == kAccCountAll) { *valid = true; return 1; }
  int64_t i = r + vals.offset[in];
  *valid = valid_bit(vals.validity[in], i);
  if (op == kAccCountValid) return *valid ? 1 : 0;
  int32_t t = vals.type[in];
  int w = vals.width[in];
  if (t == FPDB_TYPE_FLOAT64) return load_zext(vals.values[in], 8, i);
  if (t == FPDB_TYPE_FLOAT32) {
    float f = reinterpret_cast<const float*>(vals.values[in])[i];
    return (uint64_t)__double_as_longlong((double)f);
  }
  if (type_is_signed_int_d(t)) return (uint64_t)load_sext(vals.values[in], w, i);
  return load_zext(vals.values[in], w, i);
}

__device__ __forceinline__ int64_t sext_w(uint64_t raw, int width) {
  const int sh = 64 - 8 * width;
  return (int64_t)(raw << sh) >> sh;
}

// Group keys of N rows (indices clamped to valid rows by the caller), every
// key column's N values loaded together.
template <int N>
__device__ __forceinline__ void group_keys_n(const GKeySpec& k, const int64_t (&rc)[N], uint64_t (&key)[N],
                                             bool* out_of_range) {
#pragma unroll
  for (int i = 0; i < N; ++i) key[i] = 0;
#pragma unroll
  for (int j = 0; j < kMaxGroupKeys; ++j) {
    if (j >= k.n) break;
    int64_t idx[N];
    uint64_t raw[N];
#pragma unroll
    for (int i = 0; i < N; ++i) idx[i] = rc[i] + k.off[j];
    load_zext_n<N>(k.v[j], k.w[j], idx, raw);
    const uint64_t m = key_field_mask(k.bits[j]);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      int64_t v = k.sgn[j] ? sext_w(raw[i], k.w[j]) : (int64_t)raw[i];
      uint64_t d = (uint64_t)v - (uint64_t)k.bias[j];
      if (k.check && d > m) *out_of_range = true;
      key[i] |= (d & m) << k.shift[j];
    }
  }
}

// Accumulator operands of N rows (as acc_operand), each input column's N
// values (and validity bytes) loaded together.
template <int N, int NA>
__device__ __forceinline__ void acc_operands_n(const ColSet& vals, const AccSpec& A, int nacc,
                                               const int64_t (&rc)[N], uint64_t (&v)[NA][N], bool (&ok)[NA][N]) {
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    if (a >= nacc) break;
    const int32_t op = A.op[a];
    if (op == kAccCountAll) {
#pragma unroll
      for (int i = 0; i < N; ++i) { v[a][i] = 1; ok[a][i] = true; }
      continue;
    }
    const int in = A.input[a];
    int64_t idx[N];
#pragma unroll
    for (int i = 0; i < N; ++i) idx[i] = rc[i] + vals.offset[in];
    if (vals.validity[in]) {
#pragma unroll
      for (int i = 0; i < N; ++i) ok[a][i] = valid_bit(vals.validity[in], idx[i]);
    } else {
#pragma unroll
      for (int i = 0; i < N; ++i) ok[a][i] = true;
    }
    if (op == kAccCountValid) {
#pragma unroll
      for (int i = 0; i < N; ++i) v[a][i] = ok[a][i] ? 1 : 0;
      continue;
    }
    uint64_t raw[N];
    load_zext_n<N>(vals.values[in], vals.width[in], idx, raw);
    const int32_t t = vals.type[in];
    const int w = vals.width[in];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if (t == FPDB_TYPE_FLOAT64) v[a][i] = raw[i];
      else if (t == FPDB_TYPE_FLOAT32) v[a][i] = (uint64_t)__double_as_longlong((double)__uint_as_float((uint32_t)raw[i]));
      else if (type_is_signed_int_d(t)) v[a][i] = (uint64_t)sext_w(raw[i], w);
      else v[a][i] = raw[i];
    }
  }
}

__device__ __forceinline__ void atomic_min_f64(double* p, double v) {
  unsigned long long* a = reinterpret_cast<unsigned long long*>(p);
  unsigned long long old = *a;
  while (__longlong_as_double((long long)old) > v) {
    unsigned long long prev = atomicCAS(a, old, (unsigned long long)__double_as_longlong(v));
    if (prev == old) break;
    old = prev;
  }
}
__device__ __forceinline__ void atomic_max_f64(double* p, double v) {
  unsigned long long* a = reinterpret_cast<unsigned long long*>(p);
  unsigned long long old = *a;
  while (__longlong_as_double((long long)old) < v) {
    unsigned long long prev = atomicCAS(a, old, (unsigned long long)__double_as_longlong(v));
    if (prev == old) break;
    old = prev;
  }
}

// Applies one operand to an LDS accumulator.
__device__ __forceinline__ void acc_apply(uint64_t* slot, int32_t op, uint64_t v) {
  switch (op) {
    case kAccCountAll:
    case kAccCountValid:
    case kAccSumI:
      if (v) atomicAdd(reinterpret_cast<unsigned long long*>(slot), (unsigned long long)v);
      break;
    case kAccSumF:
      atomicAdd(reinterpret_cast<double*>(slot), __longlong_as_double((long long)v));
      break;
    case kAccMinI: atomicMin(reinterpret_cast<long long*>(slot), (long long)v); break;
    case kAccMaxI: atomicMax(reinterpret_cast<long long*>(slot), (long long)v); break;
    case kAccMinU: atomicMin(reinterpret_cast<unsigned long long*>(slot), (unsigned long long)v); break;
    case kAccMaxU: atomicMax(reinterpret_cast<unsigned long long*>(slot), (unsigned long long)v); break;
    case kAccMinF: atomic_min_f64(reinterpret_cast<double*>(slot), __longlong_as_double((long long)v)); break;
    case kAccMaxF: atomic_max_f64(reinterpret_cast<double*>(slot), __longlong_as_double((long long)v)); break;
  }
}

// ---- LDS hash table --------------------------------------------------------
//
// nslots ordinary slots (any count; the home slot is a multiply-shift of the
// low hash bits) plus one reserved slot, index nslots, for the key equal to
// the empty marker.  Layout: keys[stride], first[stride], acc[nacc][stride].

struct LdsTable {
  uint64_t* keys;
  long long* first;
  uint64_t* acc;
  uint32_t nslots;
  uint32_t stride;
};

__device__ __forceinline__ LdsTable lds_table(uint64_t* base, uint32_t nslots) {
  LdsTable T;
  T.nslots = nslots;
  T.stride = nslots + 1;
  T.keys = base;
  T.first = reinterpret_cast<long long*>(base + T.stride);
  T.acc = base + 2 * (size_t)T.stride;
  return T;
}

__device__ __forceinline__ size_t lds_table_words(uint32_t nslots, int nacc) {
  return (size_t)(nslots + 1) * (2 + nacc);
}

__device__ __forceinline__ uint32_t slot_home(uint64_t h, uint32_t nslots) {
  return (uint32_t)(((h & 0xFFFFFFFFULL) * (uint64_t)nslots) >> 32);
}

__device__ __forceinline__ uint32_t bucket_of(uint64_t h, int bucket_shift) {
  return bucket_shift >= 64 ? 0u : (uint32_t)(h >> bucket_shift);
}

__device__ __forceinline__ void lds_table_init(const LdsTable& T, const AccSpec& A, int nacc) {
  for (uint32_t s = threadIdx.x; s < T.stride; s += blockDim.x) {
    T.keys[s] = kEmptyKey;
    T.first[s] = LLONG_MAX;
    for (int a = 0; a < nacc; ++a) T.acc[(size_t)a * T.stride + s] = acc_identity(A.op[a]);
  }
}

// Slot of key k, claiming an empty one within max_probes of home; -1 if none.
__device__ __forceinline__ int lds_find_or_insert(const LdsTable& T, uint64_t k, uint64_t h, uint32_t max_probes) {
  if (k == kEmptyKey) return (int)T.nslots;
  uint32_t s = slot_home(h, T.nslots);
  for (uint32_t p = 0; p < max_probes; ++p) {
    uint64_t cur = T.keys[s];
    if (cur == k) return (int)s;
    if (cur == kEmptyKey) {
      unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long*>(&T.keys[s]), kEmptyKey, k);
      if (prev == kEmptyKey || prev == k) return (int)s;
    }
    s = (s + 1 == T.nslots) ? 0 : s + 1;
  }
  return -1;
}

// Lookup only (the table is no longer changing).
__device__ __forceinline__ int lds_find(const LdsTable& T, uint64_t k, uint64_t h, uint32_t max_probes) {
  if (k == kEmptyKey) return (int)T.nslots;
  uint32_t s = slot_home(h, T.nslots);
  for (uint32_t p = 0; p < max_probes; ++p) {
    uint64_t cur = T.keys[s];
    if (cur == k) return (int)s;
    if (cur == kEmptyKey) return -1;
    s = (s + 1 == T.nslots) ? 0 : s + 1;
  }
  return -1;
}

// Records of RW 8-byte words: with RW even (every record 16-byte aligned in a
// 256-byte aligned pool) they move as RW / 2 16-byte accesses, halving the
// memory instructions of the record passes.
template <int RW>
__device__ __forceinline__ void store_rec(uint64_t* rec, const uint64_t (&w)[RW]) {
  if constexpr (RW % 2 == 0) {
#pragma unroll
    for (int i = 0; i < RW; i += 2)
      reinterpret_cast<uint4*>(rec)[i / 2] = make_uint4((uint32_t)w[i], (uint32_t)(w[i] >> 32), (uint32_t)w[i + 1],
                                                        (uint32_t)(w[i + 1] >> 32));
  } else {
#pragma unroll
    for (int i = 0; i < RW; ++i) rec[i] = w[i];
  }
}

template <int RW>
__device__ __forceinline__ void load_rec(const uint64_t* rec, uint64_t (&w)[RW]) {
  if constexpr (RW % 2 == 0) {
#pragma unroll
    for (int i = 0; i < RW; i += 2) {
      const uint4 v = reinterpret_cast<const uint4*>(rec)[i / 2];
      w[i] = (uint64_t)v.x | ((uint64_t)v.y << 32);
      w[i + 1] = (uint64_t)v.z | ((uint64_t)v.w << 32);
    }
  } else {
#pragma unroll
    for (int i = 0; i < RW; ++i) w[i] = rec[i];
  }
}

// ---- kernels ---------------------------------------------------------------
//
// A consume runs as:
//   sample   (one workgroup) counts the keys of up to 65,536 sampled rows in
//            LDS and picks the most frequent ones as this batch's head keys
//            (as many as the main pass's LDS holds), each with its bucket and
//            its rank among the bucket's head keys;
//   count    (one workgroup per CU) looks every row's key up among the head
//            keys and counts the other ("tail") rows per bucket;
//   offsets  (one workgroup) turns the counts into record positions in the
//            new pool: bucket-major, old records, then head records, then
//            each workgroup's tail rows;
//   main     (same workgroups, same rows) aggregates head rows in LDS and
//            writes each tail row as a record at its position; then writes
//            its head partials;
//   merge    adds each head key's partials in workgroup order and stores its
//            record at its position;
//   compact  (one workgroup per bucket) aggregates old records, head records
//            and tail rows of the bucket into one record per group.
// A row is a head row iff its key is a head key, a fixed function of the key
// during the batch, so the count and main passes agree on every row without
// the main pass re-deriving anything from the input.  The input is read once
// in full (main) plus its key column once more (count).

constexpr int kSampleBlocks = 32;
constexpr int kSampleBlockRows = 1024;
constexpr int kHeadIndexSlots = 16384;       // head index: 16-bit head id per slot
constexpr uint32_t kNoHead = 0xFFFFFFFFu;
constexpr uint32_t kNoHead16 = 0xFFFFu;
constexpr int kCountBins = 512;
constexpr int64_t kDealRows = 4096;          // rows dealt to a workgroup at a time (count and main passes)

// Hash of a group key for its bucket (top bits) and its head-index home
// (bits 40..52): one multiply by 2^64 / golden ratio after folding the high
// half onto the low half, so every key bit reaches the bits used.
__host__ __device__ __forceinline__ uint64_t gb_hash(uint64_t k) {
  return (k ^ (k >> 31)) * 0x9E3779B97F4A7C15ULL;
}

struct GbGeom {
  int64_t n;            // rows in this batch
  int64_t row_base;     // global row index of the batch's first row
  int32_t max_heads;    // head capacity of the main pass's LDS
  int32_t nbuckets;
  int32_t bucket_shift; // 64 - log2(nbuckets)
  int32_t rec_words;    // 2 + nacc
};

// Applies the operands of ITEMS rows to their LDS slots, accumulator by
// accumulator: the operation is uniform, so its switch runs once per
// accumulator instead of once per row.
template <int ITEMS, int NA>
__device__ __forceinline__ void acc_apply_rows(const LdsTable& T, const AccSpec& A, int nacc, const int (&slot)[ITEMS],
                                               const uint64_t (&vv)[NA][ITEMS], const bool (&okk)[NA][ITEMS]) {
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    if (a >= nacc) break;
    uint64_t* acc = T.acc + (size_t)a * T.stride;
    switch (A.op[a]) {
      case kAccCountAll:
      case kAccCountValid:
      case kAccSumI:
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
          if (slot[i] >= 0 && okk[a][i] && vv[a][i])
            atomicAdd(reinterpret_cast<unsigned long long*>(acc + slot[i]), (unsigned long long)vv[a][i]);
        break;
      case kAccSumF:
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
          if (slot[i] >= 0 && okk[a][i])
            atomicAdd(reinterpret_cast<double*>(acc + slot[i]), __longlong_as_double((long long)vv[a][i]));
        break;
      default:
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
          if (slot[i] >= 0 && okk[a][i]) acc_apply(acc + slot[i], A.op[a], vv[a][i]);
        break;
    }
  }
}

// Head index: kHeadIndexSlots 16-bit head ids (0xFFFF = empty).  A head key
// sits in one of two slots, home1 or home2 of its hash, so a lookup reads two
// slots and (for occupied ones) two keys, all independent: no probe loop.  The
// sampler builds the index once per batch (a key whose two slots are both
// taken is dropped from the head set) and every workgroup of the count and
// main passes copies the same index into its LDS, so both passes classify
// every row identically.
__device__ __forceinline__ uint32_t head_home1(uint64_t h) { return (uint32_t)(h >> 40) & (kHeadIndexSlots - 1); }
__device__ __forceinline__ uint32_t head_home2(uint64_t h) { return (uint32_t)(h >> 20) & (kHeadIndexSlots - 1); }

// Head id of key k (hash h), or -1.
__device__ __forceinline__ int head_find(const uint16_t* s_idx, const uint64_t* s_hkeys, uint64_t k, uint64_t h) {
  const uint32_t a = s_idx[head_home1(h)];
  const uint32_t b = s_idx[head_home2(h)];
  const uint64_t ka = s_hkeys[a == kNoHead16 ? 0 : a];
  const uint64_t kb = s_hkeys[b == kNoHead16 ? 0 : b];
  if (a != kNoHead16 && ka == k) return (int)a;
  if (b != kNoHead16 && kb == k) return (int)b;
  return -1;
}

// Copies the batch's head index and head keys into LDS (the whole workgroup).
__device__ __forceinline__ void head_index_load(uint16_t* s_idx, uint64_t* s_hkeys, const uint16_t* __restrict__ g_idx,
                                                const uint64_t* __restrict__ hkeys, int nhead) {
  const uint4* src = reinterpret_cast<const uint4*>(g_idx);
  uint4* dst = reinterpret_cast<uint4*>(s_idx);
  for (int i = threadIdx.x; i < kHeadIndexSlots * 2 / 16; i += blockDim.x) dst[i] = src[i];
  for (int h = threadIdx.x; h < nhead; h += blockDim.x) s_hkeys[h] = hkeys[h];
  if (threadIdx.x == 0 && nhead == 0) s_hkeys[0] = 0;   // head_find reads slot 0 for empty slots
  __syncthreads();
}


// Sample, in three kernels over kSampleBlocks x kSampleBlockRows rows (all
// rows of a smaller batch; otherwise kSampleBlocks evenly spaced runs of
// consecutive rows).
//   count  (one workgroup per run) counts its run's keys in an LDS table, then
//          adds each distinct key's count to the global count table (open
//          addressing on a 64-bit CAS of the key; the key equal to the empty
//          marker counts in the reserved last slot);
//   gather (a workgroup per 1024 slots) empties the count table into a list
//          of (key, count) and a global histogram of the counts;
//   select (one workgroup) picks up to max_heads keys with the highest counts
//          (ties at the threshold count in any order), places them in the head
//          index and empties the list and histogram for the next batch.
// Outputs: the index, the head keys (ids 0 .. nhead - 1), per head key its
// rank among its bucket's head keys, per bucket its head count.  A single
// workgroup reading the whole count table would wait for memory once per
// handful of slots; the gather spreads that over the CUs.
constexpr int kSampleLocal = 2048;           // a run's LDS count table (>= 2 x kSampleBlockRows)
constexpr int kSampleTable = 1 << 16;        // global count table (>= 2 x the sampled rows)
constexpr int kSampleListMax = kSampleBlocks * kSampleBlockRows + 1;   // distinct sampled keys, at most
__global__ __launch_bounds__(kSampleBlockRows) void gb_sample_count_kernel(GKeySpec keys, GbGeom G,
                                                                           unsigned long long* __restrict__ g_keys,
                                                                           uint32_t* __restrict__ g_cnt) {
  __shared__ uint64_t s_k[kSampleLocal + 1];
  __shared__ uint32_t s_c[kSampleLocal + 1];
  const int lane = threadIdx.x & 63;
  for (int s = threadIdx.x; s <= kSampleLocal; s += blockDim.x) {
    s_k[s] = kEmptyKey;
    s_c[s] = 0;
  }
  __syncthreads();
  const int64_t n = G.n;
  const bool all = n <= (int64_t)kSampleBlocks * kSampleBlockRows;
  const int64_t j = blockIdx.x;
  const int64_t r = all ? j * kSampleBlockRows + threadIdx.x
                        : (n - kSampleBlockRows) * j / (kSampleBlocks - 1) + threadIdx.x;
  const bool live = r < n;
  int64_t rc[1] = {live ? r : n - 1};
  uint64_t kk[1];
  bool dummy = false;
  group_keys_n<1>(keys, rc, kk, &dummy);
  const uint64_t k = kk[0];
  int slot = -1;
  if (live) {
    if (k == kEmptyKey) {
      slot = kSampleLocal;
    } else {
      uint32_t s = (uint32_t)(gb_hash(k) >> 40) & (kSampleLocal - 1);
      for (int p = 0; p < kSampleLocal; ++p) {   // at most kSampleBlockRows keys: always a slot
        const unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long*>(&s_k[s]), kEmptyKey, k);
        if (prev == kEmptyKey || prev == k) { slot = (int)s; break; }
        s = (s + 1) & (kSampleLocal - 1);
      }
    }
  }
  // skewed keys repeat within a wave: twice, the lanes sharing the slot of the
  // first lane still to count add with one atomic
  bool todo = slot >= 0;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const uint64_t act = __ballot(todo);
    if (!act) break;
    const int leader = __builtin_ctzll(act);
    const int ls = __shfl(slot, leader);
    const bool same = todo && slot == ls;
    const uint64_t grp = __ballot(same);
    if (same) {
      if (lane == leader) atomicAdd(&s_c[ls], (uint32_t)__popcll(grp));
      todo = false;
    }
  }
  if (todo) atomicAdd(&s_c[slot], 1u);
  __syncthreads();
  for (int s = threadIdx.x; s <= kSampleLocal; s += blockDim.x) {
    const uint32_t c = s_c[s];
    if (!c) continue;
    uint32_t gs = kSampleTable;
    if (s < kSampleLocal) {
      const uint64_t key = s_k[s];
      gs = (uint32_t)(gb_hash(key) >> 40) & (kSampleTable - 1);
      for (int p = 0; p < kSampleTable; ++p) {   // at most the sampled rows' keys: always a slot
        const unsigned long long prev = atomicCAS(&g_keys[gs], kEmptyKey, key);
        if (prev == kEmptyKey || prev == key) break;
        gs = (gs + 1) & (kSampleTable - 1);
      }
    }
    atomicAdd(&g_cnt[gs], c);
  }
}

__global__ __launch_bounds__(1024) void gb_sample_gather_kernel(unsigned long long* __restrict__ g_keys,
                                                                uint32_t* __restrict__ g_cnt,
                                                                uint64_t* __restrict__ l_keys,
                                                                uint32_t* __restrict__ l_cnt,
                                                                uint32_t* __restrict__ l_hist) {
  __shared__ uint32_t s_hist[kCountBins];
  __shared__ uint32_t s_wcnt[16], s_base;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  for (int b = threadIdx.x; b < kCountBins; b += blockDim.x) s_hist[b] = 0;
  __syncthreads();
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;   // grid covers the slots, plus one for the reserved
  const bool in = s <= (uint32_t)kSampleTable;
  const uint32_t c = in ? g_cnt[s] : 0u;
  const uint64_t key = in && s < (uint32_t)kSampleTable ? g_keys[s] : kEmptyKey;
  if (in) {
    g_cnt[s] = 0;
    if (s < (uint32_t)kSampleTable) g_keys[s] = kEmptyKey;
  }
  if (c) atomicAdd(&s_hist[min(c, (uint32_t)(kCountBins - 1))], 1u);
  // list positions: one atomic per workgroup on the list length (kept after
  // the bins), the waves' entries in wave order after it
  const uint64_t m = __ballot(c != 0);
  if (lane == 0) s_wcnt[wave] = (uint32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += s_wcnt[w];
    s_base = t ? atomicAdd(&l_hist[kCountBins], t) : 0u;
  }
  __syncthreads();
  if (c) {
    uint32_t at = s_base + rank_below(m);
    for (int w = 0; w < wave; ++w) at += s_wcnt[w];
    l_keys[at] = key;
    l_cnt[at] = c;
  }
  for (int b = threadIdx.x; b < kCountBins; b += blockDim.x)
    if (s_hist[b]) atomicAdd(&l_hist[b], s_hist[b]);
}

__global__ __launch_bounds__(1024) void gb_sample_select_kernel(GbGeom G, const uint64_t* __restrict__ l_keys,
                                                                const uint32_t* __restrict__ l_cnt,
                                                                uint32_t* __restrict__ l_hist,
                                                                uint16_t* __restrict__ g_idx,
                                                                uint64_t* __restrict__ head_keys,
                                                                uint32_t* __restrict__ head_rank,
                                                                unsigned int* __restrict__ head_count,
                                                                int* nhead_out) {
  extern __shared__ uint64_t lds[];
  uint32_t* s_idx32 = reinterpret_cast<uint32_t*>(lds);                   // kHeadIndexSlots
  uint64_t* s_sel = lds + kHeadIndexSlots / 2;                             // max_heads selected keys
  uint32_t* s_newid = reinterpret_cast<uint32_t*>(s_sel + G.max_heads);   // max_heads
  uint32_t* s_bcount = s_newid + G.max_heads;                              // nbuckets
  __shared__ int s_thresh, s_take, s_n, s_tied, s_kept;
  const int lane = threadIdx.x & 63;
  for (int b = threadIdx.x; b < G.nbuckets; b += blockDim.x) s_bcount[b] = 0;
  for (int s = threadIdx.x; s < kHeadIndexSlots; s += blockDim.x) s_idx32[s] = kNoHead;
  if (threadIdx.x == 0) {
    s_n = 0;
    s_tied = 0;
    s_kept = 0;
    s_thresh = 1;
    s_take = 0;
  }
  const uint32_t nlist = l_hist[kCountBins];
  // threshold bin T = the highest bin with (keys in higher bins) + (keys in
  // T) >= max_heads; every key above T is taken and max_heads - (keys above)
  // of bin T.  Wave 0: lane l holds bins 8l .. 8l + 7.
  if (threadIdx.x < 64) {
    constexpr int PB = kCountBins / 64;
    uint32_t hb[PB];
    uint32_t mine = 0;
#pragma unroll
    for (int j = 0; j < PB; ++j) {
      hb[j] = l_hist[lane * PB + j];
      mine += hb[j];
    }
    uint32_t incl = mine;   // keys in bins of lanes >= lane
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t t = __shfl_down(incl, off);
      if (lane + off < 64) incl += t;
    }
    uint32_t above = incl - mine;
    int found = -1, take = 0;
#pragma unroll
    for (int j = PB - 1; j >= 0; --j) {
      const int bin = lane * PB + j;
      if (found < 0 && bin >= 1 && above + hb[j] >= (uint32_t)G.max_heads) {
        found = bin;
        take = G.max_heads - (int)above;
      }
      above += hb[j];
    }
    const uint64_t with = __ballot(found >= 0);
    if (with) {
      const int top = 63 - __clzll((long long)with);
      if (lane == top) {
        s_thresh = found;
        s_take = take;
      }
    } else if (lane == 0) {
      s_thresh = 1;                  // fewer sampled keys than head slots: take them all
      s_take = (int)hb[1];
    }
  }
  __syncthreads();
  const uint32_t T = (uint32_t)s_thresh;
  const int take_at_t = s_take;
  // a wave's selections take their places with one LDS atomic per wave (not
  // per key: thousands of keys on one counter serialise)
  auto consider = [&](uint32_t c, uint64_t key) {
    const uint32_t bin = min(c, (uint32_t)(kCountBins - 1));
    const bool tie = c != 0 && bin == T;
    bool take = c != 0 && bin > T;
    const uint64_t mt = __ballot(tie);
    if (mt) {
      uint32_t base = 0;
      if (lane == 0) base = (uint32_t)atomicAdd(&s_tied, (int)__popcll(mt));
      base = __builtin_amdgcn_readfirstlane(base);
      if (tie) take = (int)(base + rank_below(mt)) < take_at_t;
    }
    const uint64_t mk = __ballot(take);
    if (mk) {
      uint32_t base = 0;
      if (lane == 0) base = (uint32_t)atomicAdd(&s_n, (int)__popcll(mk));
      base = __builtin_amdgcn_readfirstlane(base);
      if (take) s_sel[base + rank_below(mk)] = key;
    }
  };
  // the list, kListPer entries a thread a round, their loads issued together
  constexpr int kListPer = 16;
  for (uint32_t r0 = 0; r0 < nlist; r0 += kListPer * blockDim.x) {
    uint32_t c[kListPer];
    uint64_t k[kListPer];
#pragma unroll
    for (int q = 0; q < kListPer; ++q) {
      const uint32_t i = r0 + q * blockDim.x + threadIdx.x;
      const uint32_t ic = i < nlist ? i : nlist - 1;
      c[q] = l_cnt[ic];
      k[q] = l_keys[ic];
      if (i >= nlist) c[q] = 0;
    }
#pragma unroll
    for (int q = 0; q < kListPer; ++q) consider(c[q], k[q]);
  }
  __syncthreads();
  // the list and histogram are empty again for the next batch
  for (int b = threadIdx.x; b <= kCountBins; b += blockDim.x) l_hist[b] = 0;
  const int nsel = s_n;
  for (int j = threadIdx.x; j < nsel; j += blockDim.x) {
    const uint64_t h = gb_hash(s_sel[j]);
    bool placed = atomicCAS(&s_idx32[head_home1(h)], kNoHead, (uint32_t)j) == kNoHead;
    if (!placed) placed = atomicCAS(&s_idx32[head_home2(h)], kNoHead, (uint32_t)j) == kNoHead;
    s_newid[j] = placed ? (uint32_t)atomicAdd(&s_kept, 1) : kNoHead;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < nsel; j += blockDim.x) {
    const uint32_t id = s_newid[j];
    if (id == kNoHead) continue;
    const uint64_t k = s_sel[j];
    head_keys[id] = k;
    head_rank[id] = atomicAdd(&s_bcount[bucket_of(gb_hash(k), G.bucket_shift)], 1u);
  }
  for (int s = threadIdx.x; s < kHeadIndexSlots; s += blockDim.x) {
    const uint32_t j = s_idx32[s];
    g_idx[s] = (uint16_t)(j == kNoHead ? kNoHead16 : s_newid[j]);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < G.nbuckets; b += blockDim.x) head_count[b] = s_bcount[b];
  if (threadIdx.x == 0) *nhead_out = s_kept;
}

// Rows are dealt to the count and main passes' workgroups in units of
// kDealRows consecutive rows, round-robin (unit u to workgroup u % grid), so
// all workgroups stream through neighbouring rows (HBM page locality) and both
// passes give every workgroup the same rows.
__device__ __forceinline__ int64_t gb_units(int64_t n) { return (n + kDealRows - 1) / kDealRows; }

// Count pass: tail rows per bucket of this workgroup's rows, written to
// hist[bucket][workgroup].  It runs kCountSplit workgroups per main-pass
// workgroup (its LDS allows several per CU, and more waves hide the latency of
// its key loads and index lookups); rows are dealt the same way, so count
// workgroup c sees a share of the rows of main-pass workgroup c % blocks.
// Each thread loads the keys of kCountUnits units at once (twice as many for
// a 4- or 8-byte key), 4 rows per unit; the last trip's rows past the end
// are loaded clamped, so fewer units a trip waste fewer loads (one unit a
// trip: 90 us on the 60M-row group-by workload, two: 94 us, four: 117 us).
constexpr int kCountUnits = 1;
constexpr int kCountSplit = 2;
template <int KW>
__global__ __launch_bounds__(kGroupThreads) void gb_count_kernel(GKeySpec keys, GbGeom G,
                                                                 const uint16_t* __restrict__ g_idx,
                                                                 const uint64_t* __restrict__ head_keys,
                                                                 const int* __restrict__ nhead_p,
                                                                 uint32_t* __restrict__ hist, int* key_range) {
  extern __shared__ uint64_t lds[];
  uint16_t* s_idx = reinterpret_cast<uint16_t*>(lds);
  uint64_t* s_hkeys = lds + kHeadIndexSlots / 4;
  uint32_t* s_hist = reinterpret_cast<uint32_t*>(s_hkeys + G.max_heads);
  const int nhead = *nhead_p;
  for (int b = threadIdx.x; b < G.nbuckets; b += blockDim.x) s_hist[b] = 0;
  head_index_load(s_idx, s_hkeys, g_idx, head_keys, nhead);
  bool out_of_range = false;
  constexpr int PER_UNIT = (int)(kDealRows / kGroupThreads);
  constexpr int CU = KW ? 2 * kCountUnits : kCountUnits;   // the direct key path has registers for more
  constexpr int N = CU * PER_UNIT;
  const int64_t units = gb_units(G.n);
  for (int64_t u0 = blockIdx.x; u0 < units; u0 += (int64_t)gridDim.x * CU) {
    int64_t rc[N];
    bool live[N];
#pragma unroll
    for (int j = 0; j < CU; ++j) {
#pragma unroll
      for (int i = 0; i < PER_UNIT; ++i) {
        const int64_t r = (u0 + (int64_t)j * gridDim.x) * kDealRows + (int64_t)i * kGroupThreads + threadIdx.x;
        live[j * PER_UNIT + i] = r < G.n;
        rc[j * PER_UNIT + i] = r < G.n ? r : G.n - 1;
      }
    }
    uint64_t kk[N];
    if constexpr (KW == 8) {
      const uint64_t* kp = reinterpret_cast<const uint64_t*>(keys.v[0]) + keys.off[0];
#pragma unroll
      for (int i = 0; i < N; ++i) kk[i] = kp[rc[i]];
    } else if constexpr (KW == 4) {
      const uint32_t* kp = reinterpret_cast<const uint32_t*>(keys.v[0]) + keys.off[0];
      const bool sgn = keys.sgn[0] != 0;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const uint32_t x = kp[rc[i]];
        kk[i] = sgn ? (uint64_t)(int64_t)(int32_t)x : (uint64_t)x;
      }
    } else {
      group_keys_n<N>(keys, rc, kk, &out_of_range);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const uint64_t h = gb_hash(kk[i]);
      if (live[i] && head_find(s_idx, s_hkeys, kk[i], h) < 0) atomicAdd(&s_hist[bucket_of(h, G.bucket_shift)], 1u);
    }
  }
  if (out_of_range) *key_range = 1;
  __syncthreads();
  for (int b = threadIdx.x; b < G.nbuckets; b += blockDim.x) hist[(size_t)b * gridDim.x + blockIdx.x] = s_hist[b];
}

// Offsets, first kernel (a wave per bucket): base[b][w] = tail rows of bucket
// b in main-pass workgroups before w; tail_total[b] = all of them.  The count
// pass ran kCountSplit workgroups per main-pass workgroup: hist[b][c] with
// c < kCountSplit * blocks, and count workgroup c counted rows of main-pass
// workgroup c % blocks.
__global__ __launch_bounds__(1024) void gb_bucket_scan_kernel(const uint32_t* __restrict__ hist, int blocks,
                                                              int nbuckets, uint64_t* __restrict__ base,
                                                              uint64_t* __restrict__ tail_total) {
  const int lane = threadIdx.x & 63;
  const int b = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (b >= nbuckets) return;
  const uint32_t* row = hist + (size_t)b * blocks * kCountSplit;
  constexpr int kMaxPer = 16;               // up to 1024 workgroups
  const int per = (blocks + 63) / 64;       // consecutive counts per lane
  uint32_t v[kMaxPer];
  uint64_t mine = 0;
#pragma unroll
  for (int j = 0; j < kMaxPer; ++j) {
    const int w = lane * per + j;
    uint32_t x = 0;
#pragma unroll
    for (int q = 0; q < kCountSplit; ++q) x += (j < per && w < blocks) ? row[w + q * blocks] : 0u;
    v[j] = x;
    mine += v[j];
  }
  uint64_t incl = mine;
  for (int off = 1; off < 64; off <<= 1) {
    const uint64_t t = __shfl_up(incl, off);
    if (lane >= off) incl += t;
  }
  uint64_t run = incl - mine;
#pragma unroll
  for (int j = 0; j < kMaxPer; ++j) {
    const int w = lane * per + j;
    if (j < per && w < blocks) base[(size_t)b * blocks + w] = run;
    run += v[j];
  }
  if (lane == 63) tail_total[b] = incl;
}

// Offsets, second kernel (one workgroup): for bucket b, its range of the new
// pool starts at nstart[b] and has room for its old records (ocount[b],
// aggregated in place by the compact pass), its head records (head_count[b])
// and its tail rows.  Tail rows either follow the head records in the pool
// (separate == 0: tail_start[b] = their first pool position, newcount[b] =
// head + tail records) or live in the tail pool as raw rows (separate == 1:
// tail_start[b] = their first tail-pool position, bucket-major, newcount[b] =
// head records).  Each workgroup's tail rows of b start at tail_start[b] +
// base[b][w].  totals[0] = records of the new pool; totals[2] = tail rows of
// the batch (totals[1], the group count, is summed from the buckets' record
// counts when it is read).
__global__ __launch_bounds__(1024) void gb_bucket_offsets_kernel(int nbuckets, int separate,
                                                                 const int64_t* __restrict__ ocount,
                                                                 const unsigned int* __restrict__ head_count,
                                                                 const uint64_t* __restrict__ tail_total,
                                                                 int64_t* __restrict__ nstart,
                                                                 int64_t* __restrict__ newcount,
                                                                 uint64_t* __restrict__ tail_start,
                                                                 unsigned long long* totals) {
  __shared__ uint64_t s_w[16], s_t[16];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  constexpr int kPer = kMaxBuckets / 1024;
  uint64_t rec[kPer], tl[kPer];
  uint64_t mine = 0, mine_t = 0;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int b = threadIdx.x * kPer + j;
    tl[j] = b < nbuckets ? tail_total[b] : 0;
    rec[j] = b < nbuckets ? (uint64_t)ocount[b] + head_count[b] + tl[j] : 0;
    mine += rec[j];
    mine_t += tl[j];
  }
  uint64_t incl = mine, incl_t = mine_t;
  for (int off = 1; off < 64; off <<= 1) {
    const uint64_t t = __shfl_up(incl, off);
    const uint64_t tt = __shfl_up(incl_t, off);
    if (lane >= off) {
      incl += t;
      incl_t += tt;
    }
  }
  if (lane == 63) {
    s_w[wave] = incl;
    s_t[wave] = incl_t;
  }
  __syncthreads();
  uint64_t run = incl - mine, trun = incl_t - mine_t;
  for (int w = 0; w < wave; ++w) {
    run += s_w[w];
    trun += s_t[w];
  }
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int b = threadIdx.x * kPer + j;
    if (b < nbuckets) {
      nstart[b] = (int64_t)run;
      if (separate) {
        newcount[b] = (int64_t)head_count[b];
        tail_start[b] = trun;
      } else {
        newcount[b] = (int64_t)(head_count[b] + tl[j]);
        tail_start[b] = run + (uint64_t)ocount[b] + head_count[b];
      }
    }
    run += rec[j];
    trun += tl[j];
  }
  if (threadIdx.x == blockDim.x - 1) {
    totals[0] = (unsigned long long)run;
    totals[2] = (unsigned long long)trun;
  }
}

// Main pass: head rows aggregated in LDS (slot = head id), tail rows written
// as records (key, row, operands) at their positions; then the head partials
// (first row, accumulators) to hpart[workgroup][1 + nacc][max_heads].
// Key and operand loads of ITEMS rows.  KW = 4 or 8: the key is one unpacked
// column of that width (the value itself, sign-extended when signed); V8:
// every aggregate input is an 8-byte column without nulls (its raw bits are
// the operand, COUNT operands are 1).  Otherwise the generic column walks.
// The direct loads are non-temporal: the input streams through once, and
// keeping it out of the L2 lets the tail rows' partly written lines stay there
// until they are whole (main pass 0.57 -> 0.45 ms on 60M rows).
template <int ITEMS, int NA, int KW, bool V8>
__device__ __forceinline__ void gb_load_rows(const GKeySpec& keys, const ColSet& vals, const AccSpec& A, int nacc,
                                             const int64_t (&rc)[ITEMS], uint64_t (&kk)[ITEMS],
                                             uint64_t (&vv)[NA][ITEMS], bool (&okk)[NA][ITEMS], bool* oor) {
  if constexpr (KW == 8) {
    const uint64_t* kp = reinterpret_cast<const uint64_t*>(keys.v[0]) + keys.off[0];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) kk[i] = __builtin_nontemporal_load(kp + rc[i]);
  } else if constexpr (KW == 4) {
    const uint32_t* kp = reinterpret_cast<const uint32_t*>(keys.v[0]) + keys.off[0];
    const bool sgn = keys.sgn[0] != 0;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const uint32_t x = __builtin_nontemporal_load(kp + rc[i]);
      kk[i] = sgn ? (uint64_t)(int64_t)(int32_t)x : (uint64_t)x;
    }
  } else {
    group_keys_n<ITEMS>(keys, rc, kk, oor);
  }
  if constexpr (V8) {
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      if (a >= nacc) break;
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) okk[a][i] = true;
      if (A.op[a] == kAccCountAll || A.op[a] == kAccCountValid) {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) vv[a][i] = 1;
      } else {
        const int in = A.input[a];
        const uint64_t* vp = reinterpret_cast<const uint64_t*>(vals.values[in]) + vals.offset[in];
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) vv[a][i] = __builtin_nontemporal_load(vp + rc[i]);
      }
    }
  } else {
    acc_operands_n<ITEMS, NA>(vals, A, nacc, rc, vv, okk);
  }
}

template <int NACC, int KW, bool V8>
__global__ __launch_bounds__(kGroupThreads) void gb_main_kernel(GKeySpec keys, ColSet vals, AccSpec A, GbGeom G,
                                                                const uint16_t* __restrict__ g_idx,
                                                                const uint64_t* __restrict__ head_keys,
                                                                const int* __restrict__ nhead_p,
                                                                const uint64_t* __restrict__ base,
                                                                const uint64_t* __restrict__ tail_start,
                                                                uint64_t* __restrict__ pool,
                                                                uint64_t* __restrict__ hpart) {
  extern __shared__ uint64_t lds[];
  const int nacc = NACC > 0 ? NACC : A.n;
  const int H = G.max_heads;
  uint16_t* s_idx = reinterpret_cast<uint16_t*>(lds);
  uint64_t* s_hkeys = lds + kHeadIndexSlots / 4;
  uint64_t* s_acc = s_hkeys + H;                                                     // nacc x H
  unsigned long long* s_pos = reinterpret_cast<unsigned long long*>(s_acc + (size_t)nacc * H);   // nbuckets
  uint32_t* s_first = reinterpret_cast<uint32_t*>(s_pos + G.nbuckets);              // H, batch-relative rows
  const int nhead = *nhead_p;
  for (int h = threadIdx.x; h < nhead; h += blockDim.x) {
    s_first[h] = 0xFFFFFFFFu;
    for (int a = 0; a < nacc; ++a) s_acc[(size_t)a * H + h] = acc_identity(A.op[a]);
  }
  for (int b = threadIdx.x; b < G.nbuckets; b += blockDim.x)
    s_pos[b] = tail_start[b] + base[(size_t)b * gridDim.x + blockIdx.x];
  head_index_load(s_idx, s_hkeys, g_idx, head_keys, nhead);
  LdsTable T;
  T.keys = s_hkeys;
  T.first = nullptr;
  T.acc = s_acc;
  T.nslots = (uint32_t)H;
  T.stride = (uint32_t)H;
  bool dummy = false;
  constexpr int ITEMS = NACC > 0 ? 4 : 1;
  constexpr int NA = NACC > 0 ? NACC : kMaxAggs;
  const int64_t units = gb_units(G.n);
  for (int64_t u = blockIdx.x; u < units; u += gridDim.x) {
    const int64_t ub = u * kDealRows;
    const int64_t end = ub + kDealRows < G.n ? ub + kDealRows : G.n;
    for (int64_t s0 = ub; s0 < end; s0 += (int64_t)ITEMS * kGroupThreads) {
      const int64_t r0 = s0 + threadIdx.x;
      if (r0 >= end) continue;
      int64_t rc[ITEMS];
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        const int64_t r = r0 + (int64_t)i * kGroupThreads;
        rc[i] = r < end ? r : end - 1;
      }
      uint64_t kk[ITEMS];
      uint64_t vv[NA][ITEMS];
      bool okk[NA][ITEMS];
      gb_load_rows<ITEMS, NA, KW, V8>(keys, vals, A, nacc, rc, kk, vv, okk, &dummy);
      int slot[ITEMS];
      uint64_t hh[ITEMS];
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        hh[i] = gb_hash(kk[i]);
        slot[i] = r0 + (int64_t)i * kGroupThreads < end ? head_find(s_idx, s_hkeys, kk[i], hh[i]) : -2;
      }
      acc_apply_rows<ITEMS, NA>(T, A, nacc, slot, vv, okk);
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        const uint32_t rel = (uint32_t)(r0 + (int64_t)i * kGroupThreads);
        if (slot[i] >= 0) {
          if (rel < s_first[slot[i]]) atomicMin(&s_first[slot[i]], rel);
        } else if (slot[i] == -1) {
          const uint64_t at = atomicAdd(&s_pos[bucket_of(hh[i], G.bucket_shift)], 1ULL);
          uint64_t* rec = pool + at * G.rec_words;
          if constexpr (NACC > 0) {
            uint64_t w[2 + NACC];
            w[0] = kk[i];
            w[1] = (uint64_t)(G.row_base + rel);
#pragma unroll
            for (int a = 0; a < NACC; ++a) w[2 + a] = okk[a][i] ? vv[a][i] : acc_identity(A.op[a]);
            store_rec<2 + NACC>(rec, w);
          } else {
            rec[0] = kk[i];
            rec[1] = (uint64_t)(G.row_base + rel);
#pragma unroll
            for (int a = 0; a < NA; ++a)
              if (a < nacc) rec[2 + a] = okk[a][i] ? vv[a][i] : acc_identity(A.op[a]);
          }
        }
      }
    }
  }
  __syncthreads();
  uint64_t* out = hpart + (size_t)blockIdx.x * (1 + nacc) * H;
  for (int h = threadIdx.x; h < nhead; h += blockDim.x) {
    const uint32_t f = s_first[h];
    out[h] = f == 0xFFFFFFFFu ? (uint64_t)LLONG_MAX : (uint64_t)(G.row_base + f);
    for (int a = 0; a < nacc; ++a) out[(size_t)(1 + a) * H + h] = s_acc[(size_t)a * H + h];
  }
}

// Main pass with the tail rows staged in LDS (fixed accumulator count).
// A lane-scattered record store costs the texture-address unit about as much
// as the whole wave's data whatever the number of active lanes, and only about
// a fifth of the lanes hold a tail row, so storing each tail row from the lane
// that loaded it makes the stores the bound of the pass.  Instead each wave
// appends its tail rows, in their raw tail-pool form, to a stage of its own in
// LDS, and as soon as the stage holds rpw rows (one store instruction's worth)
// the wave writes them out with every lane active, consecutive lanes writing
// consecutive parts (16 bytes, or 8 when a row has an odd number of words) of
// one row; the lane of a row's first part takes the row's tail-pool position
// (an LDS atomic on its bucket's counter) and passes it to the row's other
// lanes.  Stages are per wave, so the pass has no workgroup barriers.  A raw
// tail row is RK + nv words: the key and the batch-relative row in one word
// (RK = 1, a 4-byte key column: low half the key, high half the row) or two
// (RK = 2), then the operands of the accumulators that are not COUNT_ALL
// (those are 1).
__host__ __device__ constexpr int stage_rows_per_wave(int rr) {
  // rows a store instruction takes (64 / parts a row), plus one item's 64
  return 64 / ((rr & 1) ? rr : rr / 2) + 63;
}

template <int NACC, int KW, bool V8>
__global__ __launch_bounds__(kGroupThreads) void gb_main_staged_kernel(
    GKeySpec keys, ColSet vals, AccSpec A, GbGeom G, const uint16_t* __restrict__ g_idx,
    const uint64_t* __restrict__ head_keys, const int* __restrict__ nhead_p, const uint64_t* __restrict__ base,
    const uint64_t* __restrict__ tail_start, uint64_t* __restrict__ tpool, uint64_t* __restrict__ hpart) {
  static_assert(NACC > 0, "fixed accumulator count");
  constexpr int RK = KW == 4 ? 1 : 2;
  extern __shared__ uint4 lds_v4[];   // 16-byte aligned: the stages move 16-byte parts
  uint64_t* lds = reinterpret_cast<uint64_t*>(lds_v4);
  const int H = G.max_heads;
  uint16_t* s_idx = reinterpret_cast<uint16_t*>(lds);
  uint64_t* s_hkeys = lds + kHeadIndexSlots / 4;
  uint64_t* s_acc = s_hkeys + H;                                                     // NACC x H
  unsigned long long* s_pos = reinterpret_cast<unsigned long long*>(s_acc + (size_t)NACC * H);   // nbuckets
  // raw word of each accumulator's operand (-1: COUNT_ALL, operand 1)
  int vcol[NACC];
  int nv = 0;
#pragma unroll
  for (int a = 0; a < NACC; ++a) vcol[a] = A.op[a] == kAccCountAll ? -1 : RK + nv++;
  const int rr = RK + nv;   // words of a raw tail row
  // flush geometry: parts of pw words, ppr parts a row, rpw rows a store
  // instruction; lane / ppr by a multiply (exact for lane < 64, ppr <= 7)
  const int pw = (rr & 1) ? 1 : 2;
  const uint32_t ppr = (uint32_t)(rr / pw);
  const uint32_t magic = ppr == 1 ? 0u : 0xFFFFFFFFu / ppr + 1u;
  const uint32_t rpw = 64u / ppr;
  const int wave_rows = stage_rows_per_wave(rr);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  // stages: 16-byte aligned (H and nbuckets are even, wave_rows * rr is
  // rounded up to an even word count)
  const size_t wave_words = ((size_t)wave_rows * rr + 1) & ~(size_t)1;
  uint64_t* s_stage = reinterpret_cast<uint64_t*>(s_pos + G.nbuckets) + (size_t)wave * wave_words;
  uint32_t* s_first = reinterpret_cast<uint32_t*>(reinterpret_cast<uint64_t*>(s_pos + G.nbuckets) +
                                                  (size_t)(kGroupThreads / 64) * wave_words);   // H
  const int nhead = *nhead_p;
  for (int h = threadIdx.x; h < nhead; h += blockDim.x) {
    s_first[h] = 0xFFFFFFFFu;
#pragma unroll
    for (int a = 0; a < NACC; ++a) s_acc[(size_t)a * H + h] = acc_identity(A.op[a]);
  }
  for (int b = threadIdx.x; b < G.nbuckets; b += blockDim.x)
    s_pos[b] = tail_start[b] + base[(size_t)b * gridDim.x + blockIdx.x];
  head_index_load(s_idx, s_hkeys, g_idx, head_keys, nhead);   // ends with a barrier
  LdsTable T;
  T.keys = s_hkeys;
  T.first = nullptr;
  T.acc = s_acc;
  T.nslots = (uint32_t)H;
  T.stride = (uint32_t)H;
  bool dummy = false;
  constexpr int ITEMS = (int)(kDealRows / kGroupThreads);
  const uint32_t jl = ppr == 1 ? (uint32_t)lane : __umulhi((uint32_t)lane, magic);   // lane / ppr
  const uint32_t pp = (uint32_t)lane - jl * ppr;                                      // lane % ppr
  // writes the stage's first `rows` rows (rows <= rpw) to the tail pool
  auto flush = [&](uint32_t j0, uint32_t rows) {
    const bool live = jl < rows;
    const uint64_t* src = s_stage + (size_t)(j0 + (live ? jl : 0u)) * rr + pp * pw;
    uint64_t w0 = 0, w1 = 0;
    if (live) {
      if (pw == 2) {
        const uint4 v = *reinterpret_cast<const uint4*>(src);
        w0 = (uint64_t)v.x | ((uint64_t)v.y << 32);
        w1 = (uint64_t)v.z | ((uint64_t)v.w << 32);
      } else {
        w0 = *src;
      }
    }
    uint64_t pos = 0;
    if (live && pp == 0) {
      const uint64_t key = RK == 1 ? (KW == 4 && keys.sgn[0] ? (uint64_t)(int64_t)(int32_t)(uint32_t)w0
                                                              : (w0 & 0xFFFFFFFFULL))
                                   : w0;
      pos = atomicAdd(&s_pos[bucket_of(gb_hash(key), G.bucket_shift)], 1ULL);
    }
    pos = __shfl(pos, lane - (int)pp);   // from the lane of the row's first part
    if (live) {
      uint64_t* dst = tpool + pos * rr + pp * pw;
      if (pw == 2)
        *reinterpret_cast<uint4*>(dst) = make_uint4((uint32_t)w0, (uint32_t)(w0 >> 32), (uint32_t)w1,
                                                    (uint32_t)(w1 >> 32));
      else
        *dst = w0;
    }
  };
  uint32_t fill = 0;   // rows in this wave's stage (wave-uniform, < rpw between items)
  const int64_t units = gb_units(G.n);
  // the rows of a step are loaded while the previous step is aggregated
  uint64_t kk[ITEMS];
  uint64_t vv[NACC][ITEMS];
  bool okk[NACC][ITEMS];
  auto load_unit = [&](int64_t u, uint64_t (&k)[ITEMS], uint64_t (&v)[NACC][ITEMS], bool (&ok)[NACC][ITEMS]) {
    const int64_t ub = u * kDealRows;
    const int64_t end = ub + kDealRows < G.n ? ub + kDealRows : G.n;
    int64_t rc[ITEMS];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const int64_t r = ub + (int64_t)i * kGroupThreads + threadIdx.x;
      rc[i] = r < end ? r : end - 1;
    }
    gb_load_rows<ITEMS, NACC, KW, V8>(keys, vals, A, NACC, rc, k, v, ok, &dummy);
  };
  if ((int64_t)blockIdx.x < units) load_unit(blockIdx.x, kk, vv, okk);
  for (int64_t u = blockIdx.x; u < units; u += gridDim.x) {
    const int64_t ub = u * kDealRows;
    const int64_t end = ub + kDealRows < G.n ? ub + kDealRows : G.n;
    uint64_t nk[ITEMS];
    uint64_t nv_[NACC][ITEMS];
    bool nok[NACC][ITEMS];
    const bool more = u + gridDim.x < units;
    if (more) load_unit(u + gridDim.x, nk, nv_, nok);
    int slot[ITEMS];
    uint64_t hh[ITEMS];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      hh[i] = gb_hash(kk[i]);
      slot[i] = ub + (int64_t)i * kGroupThreads + threadIdx.x < end ? head_find(s_idx, s_hkeys, kk[i], hh[i]) : -2;
    }
    acc_apply_rows<ITEMS, NACC>(T, A, NACC, slot, vv, okk);
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const uint32_t rel = (uint32_t)(ub + (int64_t)i * kGroupThreads + threadIdx.x);
      if (slot[i] >= 0 && rel < s_first[slot[i]]) atomicMin(&s_first[slot[i]], rel);
      const uint64_t tails = __ballot(slot[i] == -1);
      if (!tails) continue;
      if (slot[i] == -1) {
        uint64_t* rec = s_stage + (size_t)(fill + rank_below(tails)) * rr;
        if (RK == 1) {
          rec[0] = (kk[i] & 0xFFFFFFFFULL) | ((uint64_t)rel << 32);
        } else {
          rec[0] = kk[i];
          rec[1] = rel;
        }
#pragma unroll
        for (int a = 0; a < NACC; ++a)
          if (vcol[a] >= 0) rec[vcol[a]] = okk[a][i] ? vv[a][i] : acc_identity(A.op[a]);
      }
      fill += (uint32_t)__popcll(tails);
      if (fill >= rpw) {
        uint32_t done = 0;
        for (; done + rpw <= fill; done += rpw) flush(done, rpw);
        // move the rest (fewer than rpw rows, all past the flushed ones) to
        // the front of the stage
        const uint32_t rest = fill - done;
        uint64_t w0 = 0, w1 = 0;
        const bool mv = jl < rest;
        const uint64_t* src = s_stage + (size_t)(done + (mv ? jl : 0u)) * rr + pp * pw;
        if (mv) {
          w0 = src[0];
          if (pw == 2) w1 = src[1];
        }
        if (mv) {
          uint64_t* dst = s_stage + (size_t)jl * rr + pp * pw;
          dst[0] = w0;
          if (pw == 2) dst[1] = w1;
        }
        fill = rest;
      }
    }
    if (more) {
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        kk[i] = nk[i];
#pragma unroll
        for (int a = 0; a < NACC; ++a) {
          vv[a][i] = nv_[a][i];
          okk[a][i] = nok[a][i];
        }
      }
    }
  }
  if (fill) flush(0, fill);
  __syncthreads();   // every head row is aggregated
  uint64_t* out = hpart + (size_t)blockIdx.x * (1 + NACC) * H;
  for (int h = threadIdx.x; h < nhead; h += blockDim.x) {
    const uint32_t f = s_first[h];
    out[h] = f == 0xFFFFFFFFu ? (uint64_t)LLONG_MAX : (uint64_t)(G.row_base + f);
#pragma unroll
    for (int a = 0; a < NACC; ++a) out[(size_t)(1 + a) * H + h] = s_acc[(size_t)a * H + h];
  }
}

// LDS bytes of the staged main pass's per-wave stages for an accumulator set
// and key form (RK = 1 for a 4-byte key column, else 2).
static size_t stage_bytes(const AccSpec& A, int rk) {
  int nv = 0;
  for (int a = 0; a < A.n; ++a) nv += A.op[a] == kAccCountAll ? 0 : 1;
  const int rr = rk + nv;
  return (size_t)(kGroupThreads / 64) * ((((size_t)stage_rows_per_wave(rr) * rr + 1) & ~(size_t)1) * 8);
}

// Applies an accumulator value to a register accumulator (merge of partials).
__device__ __forceinline__ uint64_t acc_combine(int32_t op, uint64_t x, uint64_t v) {
  switch (op) {
    case kAccCountAll:
    case kAccCountValid:
    case kAccSumI: return x + v;
    case kAccSumF: return (uint64_t)__double_as_longlong(__longlong_as_double((long long)x) +
                                                          __longlong_as_double((long long)v));
    case kAccMinI: return (long long)v < (long long)x ? v : x;
    case kAccMaxI: return (long long)v > (long long)x ? v : x;
    case kAccMinU: return v < x ? v : x;
    case kAccMaxU: return v > x ? v : x;
    case kAccMinF: return __longlong_as_double((long long)v) < __longlong_as_double((long long)x) ? v : x;
    case kAccMaxF: return __longlong_as_double((long long)v) > __longlong_as_double((long long)x) ? v : x;
    default: return x;
  }
}

// Head merge, one launch: workgroup (g, f) reduces field f (0: first row,
// 1 + a: accumulator a) of head keys [64 g, 64 g + 64) over the main pass's
// per-workgroup partials [workgroup][1 + nacc][max_heads].  Wave q takes the
// partials of workgroups [q * blocks / 16, (q + 1) * blocks / 16), all loaded
// at once, and adds them in order; wave 0 then adds the 16 sums in wave order:
// a fixed order, so head keys' FP64 sums do not depend on scheduling across
// workgroups.  Each head key's record goes to nstart[b] + ocount[b] + its
// rank among bucket b's head keys (field 0's workgroup writes the key too).
constexpr int kMergeWaves = 16;
constexpr int kMergeLoads = 16;   // partials a lane loads before adding (more go in further rounds)
__global__ __launch_bounds__(64 * kMergeWaves) void gb_head_merge_kernel(
    AccSpec A, GbGeom G, int blocks, const uint64_t* __restrict__ head_keys, const uint32_t* __restrict__ head_rank,
    const int* __restrict__ nhead_p, const uint64_t* __restrict__ hpart, const int64_t* __restrict__ nstart,
    const int64_t* __restrict__ ocount, uint64_t* __restrict__ pool) {
  __shared__ uint64_t s_sum[kMergeWaves][64];
  const int lane = threadIdx.x & 63;
  const int q = threadIdx.x >> 6;
  const int nf = 1 + A.n;
  const int f = blockIdx.y;
  const int H = G.max_heads;
  const int h = blockIdx.x * 64 + lane;
  const int nhead = *nhead_p;
  const bool live = h < nhead;
  const int hc = live ? h : 0;
  const int32_t op = f == 0 ? kAccMinI : A.op[f - 1];
  const int w0 = (int)((int64_t)blocks * q / kMergeWaves), w1 = (int)((int64_t)blocks * (q + 1) / kMergeWaves);
  const uint64_t* p = hpart + (size_t)f * H + hc;
  uint64_t x = f == 0 ? (uint64_t)LLONG_MAX : acc_identity(op);
  for (int wb = w0; wb < w1; wb += kMergeLoads) {
    uint64_t v[kMergeLoads];
#pragma unroll
    for (int j = 0; j < kMergeLoads; ++j) v[j] = p[(size_t)(wb + j < w1 ? wb + j : w1 - 1) * nf * H];
#pragma unroll
    for (int j = 0; j < kMergeLoads; ++j)
      if (wb + j < w1) x = acc_combine(op, x, v[j]);
  }
  s_sum[q][lane] = x;
  __syncthreads();
  if (q != 0 || !live) return;
  for (int r = 1; r < kMergeWaves; ++r) x = acc_combine(op, x, s_sum[r][lane]);
  const uint64_t k = head_keys[h];
  const uint32_t b = bucket_of(gb_hash(k), G.bucket_shift);
  uint64_t* rec = pool + (size_t)(nstart[b] + ocount[b] + head_rank[h]) * G.rec_words;
  if (f == 0) rec[0] = k;
  rec[1 + f] = x;
}

// Compact pass, one workgroup per bucket: old records (previous batches, one
// per group) and the new ones are aggregated in LDS and written back, one per
// group, at the start of the bucket's range in the new pool.  The new records
// are the head records in the pool and the tail rows: after the head records
// in the pool (tpool == nullptr) or raw in the tail pool (rr words each, key
// and batch-relative row packed as the staged main pass writes them: rk = 1 a
// 4-byte key zero-extended, rk = 3 a 4-byte key sign-extended, rk = 2 key and
// row in two words).
template <int NACC>
__global__ __launch_bounds__(kGroupThreads) void gb_compact_kernel(AccSpec A, uint32_t nslots, int rec_words,
                                                                   const uint64_t* __restrict__ oldpool,
                                                                   int64_t* __restrict__ ostart,
                                                                   int64_t* __restrict__ ocount,
                                                                   uint64_t* pool, const int64_t* __restrict__ nstart,
                                                                   const int64_t* __restrict__ newcount,
                                                                   const uint64_t* __restrict__ tpool,
                                                                   const uint64_t* __restrict__ tail_start,
                                                                   const uint64_t* __restrict__ tail_total, int rk,
                                                                   int rr, int64_t row_base,
                                                                   int* overflow) {
  extern __shared__ uint64_t lds[];
  __shared__ uint32_t s_n;
  __shared__ int s_over;
  const int nacc = NACC > 0 ? NACC : A.n;
  const int p = blockIdx.x;
  LdsTable T = lds_table(lds, nslots);
  lds_table_init(T, A, nacc);
  if (threadIdx.x == 0) {
    s_n = 0;
    s_over = 0;
  }
  __syncthreads();
  const int64_t nold = ocount[p], nnew = newcount[p];
  const int64_t ntail = tpool ? (int64_t)tail_total[p] : 0;
  const int64_t start = nstart[p];
  const uint64_t* news = pool + (size_t)(start + nold) * rec_words;
  const uint64_t* tails = tpool ? tpool + (size_t)tail_start[p] * rr : nullptr;
  constexpr int RW = NACC > 0 ? 2 + NACC : 2 + kMaxAggs;
  // raw operand word of each accumulator (-1: COUNT_ALL, operand 1)
  int vcol[NACC > 0 ? NACC : 1];
  if constexpr (NACC > 0) {
    int nv = 0;
#pragma unroll
    for (int a = 0; a < NACC; ++a) vcol[a] = A.op[a] == kAccCountAll ? -1 : (rk == 2 ? 2 : 1) + nv++;
  }
  // R records per thread per trip, all loaded before any is aggregated
  constexpr int R = NACC > 0 ? 4 : 1;
  const int64_t nrec = nold + nnew;
  auto apply = [&](const uint64_t* w) {
    const uint64_t k = w[0];
    const int s = lds_find_or_insert(T, k, fmix64(k), nslots);
    if (s < 0) {
      s_over = 1;
      return;
    }
#pragma unroll
    for (int a = 0; a < (NACC > 0 ? NACC : kMaxAggs); ++a) {
      if (a < nacc) {
        const uint64_t v = w[2 + a];
        if (v != acc_identity(A.op[a]) || A.op[a] == kAccSumF) acc_apply(&T.acc[(size_t)a * T.stride + s], A.op[a], v);
      }
    }
    const long long f = (long long)w[1];
    if (f < T.first[s]) atomicMin(&T.first[s], f);
  };
  // records: old ones, then the new ones in the pool
  for (int64_t t0 = 0; t0 < nrec; t0 += (int64_t)R * blockDim.x) {
    const int64_t i0 = t0 + threadIdx.x;
    uint64_t w[R][RW];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      int64_t i = i0 + (int64_t)j * blockDim.x;
      i = i < nrec ? i : nrec - 1;
      const uint64_t* rec = i < nold ? oldpool + (size_t)(ostart[p] + i) * rec_words
                                     : news + (size_t)(i - nold) * rec_words;
      if constexpr (NACC > 0) {
        load_rec<RW>(rec, w[j]);
      } else {
        for (int x = 0; x < rec_words; ++x) w[j][x] = rec[x];
      }
    }
#pragma unroll
    for (int j = 0; j < R; ++j)
      if (i0 + (int64_t)j * blockDim.x < nrec) apply(w[j]);
  }
  // raw tail rows (every row's words loaded before any is converted; words
  // are loaded by their runtime position, not picked from a register array,
  // which would put the array in scratch memory)
  if constexpr (NACC > 0) {
    for (int64_t t0 = 0; t0 < ntail; t0 += (int64_t)R * blockDim.x) {
      const int64_t i0 = t0 + threadIdx.x;
      uint64_t w[R][RW];
#pragma unroll
      for (int j = 0; j < R; ++j) {
        int64_t i = i0 + (int64_t)j * blockDim.x;
        i = i < ntail ? i : ntail - 1;
        const uint64_t* raw = tails + (size_t)i * rr;
        w[j][0] = raw[0];
        w[j][1] = rk == 2 ? raw[1] : 0;
#pragma unroll
        for (int a = 0; a < NACC; ++a) w[j][2 + a] = vcol[a] < 0 ? 1ULL : raw[vcol[a]];
      }
#pragma unroll
      for (int j = 0; j < R; ++j) {
        if (i0 + (int64_t)j * blockDim.x >= ntail) continue;
        if (rk == 2) {
          w[j][1] = (uint64_t)(row_base + (int64_t)(uint32_t)w[j][1]);
        } else {
          const uint64_t x0 = w[j][0];
          const uint32_t k32 = (uint32_t)x0;
          w[j][0] = rk == 3 ? (uint64_t)(int64_t)(int32_t)k32 : (uint64_t)k32;
          w[j][1] = (uint64_t)(row_base + (int64_t)(x0 >> 32));
        }
        apply(w[j]);
      }
    }
  }
  __syncthreads();   // every record of the bucket is read before any is overwritten
  uint64_t* out = pool + (size_t)start * rec_words;
  for (uint32_t s = threadIdx.x; s < T.stride; s += blockDim.x) {
    long long f = T.first[s];
    if (f == LLONG_MAX) continue;
    uint32_t idx = atomicAdd(&s_n, 1u);
    uint64_t* rec = out + (size_t)idx * rec_words;
    rec[0] = s == T.nslots ? kEmptyKey : T.keys[s];
    rec[1] = (uint64_t)f;
    for (int a = 0; a < nacc; ++a) rec[2 + a] = T.acc[(size_t)a * T.stride + s];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ostart[p] = start;
    ocount[p] = s_n;
    if (s_over) *overflow = 1;
  }
}

// Balanced compact, for the staged main pass's layout (raw tail rows in the
// tail pool).  The buckets' items -- old records, head records, raw tail rows,
// in that order, bucket after bucket -- are numbered by record position:
// bucket b's items are the positions [nstart[b], nstart[b] + its item count)
// of the new pool, all of them [0, total).  Piece q (one workgroup per CU)
// takes positions [total * q / P, total * (q + 1) / P), so every CU streams
// the same number of items whatever the bucket sizes (with one workgroup per
// bucket the largest bucket set the time).  For each bucket it overlaps, a
// piece aggregates the overlapping items in an LDS table and writes one
// record per group at the segment's first position (a segment of m items has
// at most m groups, so it never writes past its own positions).  A bucket
// inside one piece is then final (its record count in bcount[b]); a bucket
// split over pieces q0..q1 has one partial record set per piece, counted in
// seg_last[q0], seg_first[q] = seg_last[q] for q0 < q < q1, and seg_first[q1],
// which gb_compact_merge_kernel combines.
__device__ __forceinline__ int64_t piece_lo(int64_t total, int q, int pieces) {
  return (int64_t)((uint64_t)total * (uint64_t)q / (uint64_t)pieces);   // total < 2^40, q <= 1024
}

__device__ __forceinline__ int piece_of(int64_t x, int64_t total, int pieces) {
  int q = (int)((uint64_t)x * (uint64_t)pieces / (uint64_t)total);
  if (q >= pieces) q = pieces - 1;
  while (q + 1 < pieces && piece_lo(total, q + 1, pieces) <= x) ++q;
  while (q > 0 && piece_lo(total, q, pieces) > x) --q;
  return q;
}

template <int NACC>
__global__ __launch_bounds__(kGroupThreads) void gb_compact_piece_kernel(
    AccSpec A, uint32_t nslots, int rec_words, int nbuckets, const unsigned long long* __restrict__ totals,
    const uint64_t* __restrict__ oldpool, const int64_t* __restrict__ ostart, const int64_t* __restrict__ ocount,
    uint64_t* pool, const int64_t* __restrict__ nstart, const int64_t* __restrict__ newcount,
    const uint64_t* __restrict__ tpool, const uint64_t* __restrict__ tail_start,
    const uint64_t* __restrict__ tail_total, int rk, int rr, int64_t row_base, int64_t* __restrict__ bcount,
    int64_t* __restrict__ seg_first, int64_t* __restrict__ seg_last, int* overflow) {
  static_assert(NACC > 0, "fixed accumulator count");
  extern __shared__ uint64_t lds[];
  __shared__ uint32_t s_n;
  __shared__ int s_over;
  const int pieces = gridDim.x, q = blockIdx.x;
  const int64_t total = (int64_t)totals[0];
  const int64_t lo = piece_lo(total, q, pieces), hi = piece_lo(total, q + 1, pieces);
  if (lo >= hi) return;
  // the bucket holding position lo: the last one starting at or before it
  int b = 0;
  for (int b1 = nbuckets - 1; b < b1;) {
    const int m = (b + b1 + 1) >> 1;
    if (nstart[m] <= lo) b = m;
    else b1 = m - 1;
  }
  LdsTable T = lds_table(lds, nslots);
  constexpr int RW = 2 + NACC;
  int vcol[NACC];   // raw operand word of each accumulator (-1: COUNT_ALL, operand 1)
  {
    int nv = 0;
#pragma unroll
    for (int a = 0; a < NACC; ++a) vcol[a] = A.op[a] == kAccCountAll ? -1 : (rk == 2 ? 2 : 1) + nv++;
  }
  constexpr int R = 4;   // items per thread per trip, all loaded before any is aggregated
  if (threadIdx.x == 0) s_over = 0;
  auto apply = [&](const uint64_t* w) {
    const uint64_t k = w[0];
    const int s = lds_find_or_insert(T, k, fmix64(k), nslots);
    if (s < 0) {
      s_over = 1;
      return;
    }
#pragma unroll
    for (int a = 0; a < NACC; ++a) {
      const uint64_t v = w[2 + a];
      if (v != acc_identity(A.op[a]) || A.op[a] == kAccSumF) acc_apply(&T.acc[(size_t)a * T.stride + s], A.op[a], v);
    }
    const long long f = (long long)w[1];
    if (f < T.first[s]) atomicMin(&T.first[s], f);
  };
  for (; b < nbuckets && nstart[b] < hi; ++b) {
    const int64_t nold = ocount[b], nnew = newcount[b], ntail = (int64_t)tail_total[b];
    const int64_t bs = nstart[b], be = bs + nold + nnew + ntail;
    const int64_t s0 = lo > bs ? lo : bs, s1 = hi < be ? hi : be;
    if (s0 >= s1) continue;
    lds_table_init(T, A, NACC);
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    // records: old ones (previous batches), then the head records in the pool
    const int64_t nrec = nold + nnew;
    const int64_t r0 = s0 - bs, r1 = s1 - bs < nrec ? s1 - bs : nrec;
    for (int64_t t0 = r0; t0 < r1; t0 += (int64_t)R * blockDim.x) {
      const int64_t i0 = t0 + threadIdx.x;
      uint64_t w[R][RW];
#pragma unroll
      for (int j = 0; j < R; ++j) {
        int64_t i = i0 + (int64_t)j * blockDim.x;
        i = i < r1 ? i : r1 - 1;
        const uint64_t* rec = i < nold ? oldpool + (size_t)(ostart[b] + i) * rec_words : pool + (size_t)(bs + i) * rec_words;
        load_rec<RW>(rec, w[j]);
      }
#pragma unroll
      for (int j = 0; j < R; ++j)
        if (i0 + (int64_t)j * blockDim.x < r1) apply(w[j]);
    }
    // raw tail rows (words loaded by their runtime position, not picked from
    // a register array, which would put the array in scratch memory), RT rows
    // a thread per trip, the next trip's rows loaded before this trip's are
    // aggregated (two trips' rows in registers)
    const int64_t j0 = s0 - bs - nrec > 0 ? s0 - bs - nrec : 0, j1 = s1 - bs - nrec;
    const uint64_t* tails = tpool + (size_t)tail_start[b] * rr;
    constexpr int RT = 2;
    auto load_trip = [&](int64_t t0, uint64_t (&w)[RT][RW]) {
#pragma unroll
      for (int j = 0; j < RT; ++j) {
        int64_t i = t0 + threadIdx.x + (int64_t)j * blockDim.x;
        i = i < j1 ? i : j1 - 1;
        const uint64_t* raw = tails + (size_t)i * rr;
        w[j][0] = raw[0];
        w[j][1] = rk == 2 ? raw[1] : 0;
#pragma unroll
        for (int a = 0; a < NACC; ++a) w[j][2 + a] = vcol[a] < 0 ? 1ULL : raw[vcol[a]];
      }
    };
    uint64_t nw[RT][RW];
    if (j0 < j1) load_trip(j0, nw);
    for (int64_t t0 = j0; t0 < j1; t0 += (int64_t)RT * blockDim.x) {
      const int64_t i0 = t0 + threadIdx.x;
      uint64_t w[RT][RW];
#pragma unroll
      for (int j = 0; j < RT; ++j)
#pragma unroll
        for (int x = 0; x < RW; ++x) w[j][x] = nw[j][x];
      // past the last trip this reloads the current one (unused)
      load_trip(t0 + (int64_t)RT * blockDim.x < j1 ? t0 + (int64_t)RT * blockDim.x : t0, nw);
#pragma unroll
      for (int j = 0; j < RT; ++j) {
        if (i0 + (int64_t)j * blockDim.x >= j1) continue;
        if (rk == 2) {
          w[j][1] = (uint64_t)(row_base + (int64_t)(uint32_t)w[j][1]);
        } else {
          const uint64_t x0 = w[j][0];
          const uint32_t k32 = (uint32_t)x0;
          w[j][0] = rk == 3 ? (uint64_t)(int64_t)(int32_t)k32 : (uint64_t)k32;
          w[j][1] = (uint64_t)(row_base + (int64_t)(x0 >> 32));
        }
        apply(w[j]);
      }
    }
    __syncthreads();   // every item of the segment is read before any record is written
    uint64_t* out = pool + (size_t)s0 * rec_words;
    for (uint32_t s = threadIdx.x; s < T.stride; s += blockDim.x) {
      const long long f = T.first[s];
      if (f == LLONG_MAX) continue;
      const uint32_t idx = atomicAdd(&s_n, 1u);
      uint64_t* rec = out + (size_t)idx * rec_words;
      rec[0] = s == T.nslots ? kEmptyKey : T.keys[s];
      rec[1] = (uint64_t)f;
      for (int a = 0; a < NACC; ++a) rec[2 + a] = T.acc[(size_t)a * T.stride + s];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const bool before = bs < lo, after = be > hi;
      if (!before && !after) bcount[b] = s_n;
      if (before) seg_first[q] = s_n;
      if (after) seg_last[q] = s_n;
    }
  }
  if (threadIdx.x == 0 && s_over) *overflow = 1;
}

// After the pieces, one workgroup per bucket: a bucket inside one piece gets
// its count; a split one has its partial record sets aggregated in LDS and
// written back, one record per group, at the start of its range.
template <int NACC>
__global__ __launch_bounds__(kGroupThreads) void gb_compact_merge_kernel(
    AccSpec A, uint32_t nslots, int rec_words, int pieces, const unsigned long long* __restrict__ totals,
    uint64_t* pool, const int64_t* __restrict__ nstart, int64_t* __restrict__ ostart, int64_t* __restrict__ ocount,
    const int64_t* __restrict__ newcount, const uint64_t* __restrict__ tail_total,
    const int64_t* __restrict__ bcount, const int64_t* __restrict__ seg_first,
    const int64_t* __restrict__ seg_last, int* overflow) {
  static_assert(NACC > 0, "fixed accumulator count");
  extern __shared__ uint64_t lds[];
  __shared__ uint32_t s_n;
  __shared__ int s_over;
  const int b = blockIdx.x;
  const int64_t bs = nstart[b];
  const int64_t items = ocount[b] + newcount[b] + (int64_t)tail_total[b];
  const int64_t total = (int64_t)totals[0];
  const int q0 = items ? piece_of(bs, total, pieces) : 0, q1 = items ? piece_of(bs + items - 1, total, pieces) : 0;
  if (items == 0 || q0 == q1) {
    if (threadIdx.x == 0) {
      ostart[b] = bs;
      ocount[b] = items ? bcount[b] : 0;
    }
    return;
  }
  LdsTable T = lds_table(lds, nslots);
  lds_table_init(T, A, NACC);
  if (threadIdx.x == 0) {
    s_n = 0;
    s_over = 0;
  }
  __syncthreads();
  constexpr int RW = 2 + NACC;
  for (int q = q0; q <= q1; ++q) {
    const int64_t lo = piece_lo(total, q, pieces);
    if (lo == piece_lo(total, q + 1, pieces)) continue;   // an empty piece (fewer items than pieces) wrote nothing
    const int64_t s0 = lo > bs ? lo : bs;
    const int64_t cnt = q == q0 ? seg_last[q] : seg_first[q];
    for (int64_t i = threadIdx.x; i < cnt; i += blockDim.x) {
      uint64_t w[RW];
      load_rec<RW>(pool + (size_t)(s0 + i) * rec_words, w);
      const uint64_t k = w[0];
      const int s = lds_find_or_insert(T, k, fmix64(k), nslots);
      if (s < 0) {
        s_over = 1;
        continue;
      }
#pragma unroll
      for (int a = 0; a < NACC; ++a) {
        const uint64_t v = w[2 + a];
        if (v != acc_identity(A.op[a]) || A.op[a] == kAccSumF) acc_apply(&T.acc[(size_t)a * T.stride + s], A.op[a], v);
      }
      const long long f = (long long)w[1];
      if (f < T.first[s]) atomicMin(&T.first[s], f);
    }
  }
  __syncthreads();   // every partial record is read before any is overwritten
  uint64_t* out = pool + (size_t)bs * rec_words;
  for (uint32_t s = threadIdx.x; s < T.stride; s += blockDim.x) {
    const long long f = T.first[s];
    if (f == LLONG_MAX) continue;
    const uint32_t idx = atomicAdd(&s_n, 1u);
    uint64_t* rec = out + (size_t)idx * rec_words;
    rec[0] = s == T.nslots ? kEmptyKey : T.keys[s];
    rec[1] = (uint64_t)f;
    for (int a = 0; a < NACC; ++a) rec[2 + a] = T.acc[(size_t)a * T.stride + s];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ostart[b] = bs;
    ocount[b] = s_n;
    if (s_over) *overflow = 1;
  }
}

// Groups so far: the sum of the buckets' record counts into *out (one
// workgroup; the compact passes keep per-bucket counts only, so that their
// workgroups do not queue on one global counter as they finish together).
__global__ __launch_bounds__(1024) void gb_sum_groups_kernel(const int64_t* __restrict__ ocount, int nbuckets,
                                                             unsigned long long* __restrict__ out) {
  __shared__ unsigned long long s_w[16];
  unsigned long long v = 0;
  for (int b = threadIdx.x; b < nbuckets; b += blockDim.x) v += (unsigned long long)ocount[b];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += s_w[w];
    *out = t;
  }
}

// ---- finalize ----------------------------------------------------------------

// Marks the first row of every group.
__global__ void gb_mark_kernel(const uint64_t* __restrict__ pool, const int64_t* __restrict__ ostart,
                               const int64_t* __restrict__ ocount, int rec_words, unsigned long long* bits) {
  const int p = blockIdx.x;
  const int64_t n = ocount[p];
  const uint64_t* recs = pool + (size_t)ostart[p] * rec_words;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    int64_t f = (int64_t)recs[(size_t)i * rec_words + 1];
    atomicOr(&bits[f >> 6], 1ULL << (f & 63));
  }
}

// Word-level exclusive popcount prefix: chunk of 1024 words per block.
constexpr int kScanChunk = 1024;
__global__ __launch_bounds__(1024) void popcount_chunk_kernel(const unsigned long long* __restrict__ bits,
                                                              int64_t words, uint32_t* __restrict__ chunk_sums) {
  __shared__ uint32_t s_w[16];
  int64_t i = (int64_t)blockIdx.x * kScanChunk + threadIdx.x;
  uint32_t v = i < words ? (uint32_t)__popcll(bits[i]) : 0u;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < 16; ++w) t += s_w[w];
    chunk_sums[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(1024) void chunk_scan_kernel(uint32_t* __restrict__ chunk_sums, int64_t chunks) {
  // single block, sequential over groups of 1024
  __shared__ uint32_t s_w[16];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < chunks; base += 1024) {
    int64_t i = base + threadIdx.x;
    uint32_t v = i < chunks ? chunk_sums[i] : 0u;
    uint32_t incl = wave_incl_scan(v);
    if ((threadIdx.x & 63) == 63) s_w[threadIdx.x >> 6] = incl;
    __syncthreads();
    uint32_t wprefix = 0;
    for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) wprefix += s_w[w];
    uint32_t carry = s_carry;
    if (i < chunks) chunk_sums[i] = carry + wprefix + incl - v;
    __syncthreads();
    if (threadIdx.x == 1023) s_carry = carry + wprefix + incl;
    __syncthreads();
  }
}

__global__ __launch_bounds__(1024) void word_prefix_kernel(const unsigned long long* __restrict__ bits,
                                                           int64_t words, const uint32_t* __restrict__ chunk_prefix,
                                                           uint32_t* __restrict__ word_prefix) {
  __shared__ uint32_t s_w[16];
  int64_t i = (int64_t)blockIdx.x * kScanChunk + threadIdx.x;
  uint32_t v = i < words ? (uint32_t)__popcll(bits[i]) : 0u;
  uint32_t incl = wave_incl_scan(v);
  if ((threadIdx.x & 63) == 63) s_w[threadIdx.x >> 6] = incl;
  __syncthreads();
  uint32_t wprefix = 0;
  for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) wprefix += s_w[w];
  if (i < words) word_prefix[i] = chunk_prefix[blockIdx.x] + wprefix + incl - v;
}

struct GOutSpec {
  uint8_t* key_values[kMaxGroupKeys];
  int32_t key_w[kMaxGroupKeys];
  int64_t key_off[kMaxGroupKeys];
  int64_t key_bias[kMaxGroupKeys];
  int32_t key_bits[kMaxGroupKeys];
  int32_t key_shift[kMaxGroupKeys];
  uint8_t* agg_values[kMaxAggs];
  uint8_t* agg_validity[kMaxAggs];
  int64_t agg_off[kMaxAggs];
  int32_t agg_w[kMaxAggs];
  int32_t agg_count_acc[kMaxAggs];   // hidden valid-count accumulator, -1 if none
  int32_t agg_is_f32[kMaxAggs];      // MIN/MAX of float32: narrow back
  int32_t agg_mean[kMaxAggs];        // 0: not a mean, 1: integer sum / count, 2: float sum / count
  int32_t num_keys, num_aggs;
};

// One workgroup per bucket: every group record to its first-appearance rank.
__global__ void gb_emit_kernel(const uint64_t* __restrict__ pool, const int64_t* __restrict__ ostart,
                               const int64_t* __restrict__ ocount, int rec_words,
                               const unsigned long long* __restrict__ bits,
                               const uint32_t* __restrict__ word_prefix, GOutSpec o) {
  const int p = blockIdx.x;
  const int64_t n = ocount[p];
  const uint64_t* recs = pool + (size_t)ostart[p] * rec_words;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const uint64_t* rec = recs + (size_t)i * rec_words;
    int64_t f = (int64_t)rec[1];
    uint64_t w = bits[f >> 6];
    uint64_t below = (f & 63) ? (w & (~0ULL >> (64 - (f & 63)))) : 0ULL;
    int64_t rank = (int64_t)word_prefix[f >> 6] + __popcll(below);
    uint64_t key = rec[0];
    for (int j = 0; j < o.num_keys; ++j) {
      uint64_t v = ((key >> o.key_shift[j]) & key_field_mask(o.key_bits[j])) + (uint64_t)o.key_bias[j];
      store_w(o.key_values[j], o.key_w[j], rank + o.key_off[j], v);
    }
    for (int a = 0; a < o.num_aggs; ++a) {
      uint64_t v = rec[2 + a];
      if (o.agg_is_f32[a]) {
        float fv = (float)__longlong_as_double((long long)v);
        v = (uint64_t)__float_as_uint(fv);
      }
      uint64_t cnt = o.agg_count_acc[a] >= 0 ? rec[2 + o.agg_count_acc[a]] : 1;
      if (o.agg_mean[a] && cnt) {
        double mean = o.agg_mean[a] == 1 ? (double)((long long)v / (long long)cnt)
                                         : __longlong_as_double((long long)v) / (double)cnt;
        v = (uint64_t)__double_as_longlong(mean);
      }
      int64_t di = rank + o.agg_off[a];
      store_w(o.agg_values[a], o.agg_w[a], di, v);
      bool valid = cnt != 0;
      if (o.agg_validity[a] && valid)
        atomicOr(reinterpret_cast<unsigned int*>(o.agg_validity[a]) + (di >> 5), 1u << (di & 31));
    }
  }
}

// Minimum and maximum of the valid values of an integer column, in an
// order-preserving unsigned encoding (signed values with the sign bit flipped)
// so one unsigned min/max serves both.  Each workgroup stores its partial
// (min, max, count) at part[b], part[B + b], part[2B + b]; one workgroup then
// folds them into out = [min, max, valid count].
__global__ __launch_bounds__(1024) void range_fold_kernel(const unsigned long long* __restrict__ part, int blocks,
                                                          unsigned long long* out) {
  __shared__ unsigned long long s_lo[16], s_hi[16], s_c[16];
  unsigned long long lo = ~0ULL, hi = 0ULL, cnt = 0;
  for (int b = threadIdx.x; b < blocks; b += blockDim.x) {
    if (!part[2 * blocks + b]) continue;
    lo = part[b] < lo ? part[b] : lo;
    hi = part[blocks + b] > hi ? part[blocks + b] : hi;
    cnt += part[2 * blocks + b];
  }
  for (int d = 32; d >= 1; d >>= 1) {
    unsigned long long olo = __shfl_xor(lo, d), ohi = __shfl_xor(hi, d), oc = __shfl_xor(cnt, d);
    lo = olo < lo ? olo : lo;
    hi = ohi > hi ? ohi : hi;
    cnt += oc;
  }
  if ((threadIdx.x & 63) == 0) {
    s_lo[threadIdx.x >> 6] = lo;
    s_hi[threadIdx.x >> 6] = hi;
    s_c[threadIdx.x >> 6] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      lo = s_lo[w] < lo ? s_lo[w] : lo;
      hi = s_hi[w] > hi ? s_hi[w] : hi;
      cnt += s_c[w];
    }
    out[0] = lo;
    out[1] = hi;
    out[2] = cnt;
  }
}

__global__ __launch_bounds__(256) void range_kernel(ColSet col, int64_t n, unsigned long long* part) {
  __shared__ unsigned long long s_lo[4], s_hi[4], s_c[4];
  const bool sgn = type_is_signed_int_d(col.type[0]);
  unsigned long long lo = ~0ULL, hi = 0ULL, cnt = 0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = r + col.offset[0];
    if (!valid_bit(col.validity[0], i)) continue;
    unsigned long long u = sgn ? ((unsigned long long)load_sext(col.values[0], col.width[0], i) ^ (1ULL << 63))
                               : (unsigned long long)load_zext(col.values[0], col.width[0], i);
    lo = u < lo ? u : lo;
    hi = u > hi ? u : hi;
    ++cnt;
  }
  for (int d = 32; d >= 1; d >>= 1) {
    unsigned long long olo = __shfl_xor(lo, d), ohi = __shfl_xor(hi, d), oc = __shfl_xor(cnt, d);
    lo = olo < lo ? olo : lo;
    hi = ohi > hi ? ohi : hi;
    cnt += oc;
  }
  if ((threadIdx.x & 63) == 0) {
    s_lo[threadIdx.x >> 6] = lo;
    s_hi[threadIdx.x >> 6] = hi;
    s_c[threadIdx.x >> 6] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) {
      lo = s_lo[w] < lo ? s_lo[w] : lo;
      hi = s_hi[w] > hi ? s_hi[w] : hi;
      cnt += s_c[w];
    }
    part[blockIdx.x] = lo;
    part[gridDim.x + blockIdx.x] = hi;
    part[2 * gridDim.x + blockIdx.x] = cnt;
  }
}

}  // namespace fpdb

struct fpdb_group_table {
  fpdb::Context* ctx = nullptr;
  int num_keys = 0;
  int32_t key_types[fpdb::kMaxGroupKeys] = {0, 0, 0, 0};
  int64_t key_bias[fpdb::kMaxGroupKeys] = {0, 0, 0, 0};
  int32_t key_bits[fpdb::kMaxGroupKeys] = {0, 0, 0, 0};
  int32_t key_shift[fpdb::kMaxGroupKeys] = {0, 0, 0, 0};
  int key_check = 0;                 // packed keys: values outside their ranges are an error
  int num_aggs = 0;
  fpdb_agg_spec aggs[fpdb::kMaxAggs];
  fpdb::AccSpec acc;                 // visible aggregates first, then hidden valid counts
  int32_t hidden_of[fpdb::kMaxAggs];  // accumulator index of agg a's valid count, -1 if none
  int64_t max_groups = 0;
  int rec_words = 0;                 // 2 + accumulators
  int nbuckets = 1, bucket_shift = 64;
  int max_heads = 0;                 // head keys the main pass's LDS holds
  uint32_t compact_slots = 0;        // LDS slots of the compact pass
  size_t sample_lds = 0, count_lds = 0, main_lds = 0, compact_lds = 0;
  int max_blocks = 0;                // workgroups of the count and main passes (one per CU)
  int64_t pool_records_bound = 0;    // records the pool can hold after the last consume (an upper bound)
  // device buffers
  uint64_t* pool = nullptr;          // one record per group so far (bucket ranges ostart/ocount)
  int64_t pool_cap = 0;              // records
  uint64_t* spare = nullptr;         // the next pool
  int64_t spare_cap = 0;
  unsigned long long* sample_keys = nullptr;   // the sampler's count table (kSampleTable + 1), empty between batches
  uint32_t* sample_cnt = nullptr;
  uint64_t* list_keys = nullptr;     // the sampled keys with their counts, and the counts' histogram
  uint32_t* list_cnt = nullptr;      // (kCountBins bins, then the list length); empty between batches
  uint32_t* list_hist = nullptr;
  uint64_t* tpool = nullptr;         // raw tail rows of the current batch (staged main pass)
  int64_t tpool_words = 0;
  int64_t* ostart = nullptr;         // nbuckets
  int64_t* ocount = nullptr;         // nbuckets
  int64_t* nstart = nullptr;         // nbuckets
  int64_t* newcount = nullptr;       // nbuckets
  uint64_t* base = nullptr;          // nbuckets x max_blocks tail positions, bucket-major
  uint32_t* hist = nullptr;          // nbuckets x (kCountSplit x max_blocks) tail counts, bucket-major
  uint16_t* head_index = nullptr;    // kHeadIndexSlots head ids (two-choice index of the head keys)
  uint64_t* head_keys = nullptr;     // max_heads
  uint32_t* head_rank = nullptr;     // max_heads: rank among its bucket's head keys
  unsigned int* head_count = nullptr;   // nbuckets
  int* nhead = nullptr;              // head keys of the current batch
  uint64_t* hpart = nullptr;         // max_blocks x (1 + accumulators) x max_heads head partials
  int64_t* bcount = nullptr;         // nbuckets: records of a bucket the compact finished in one piece
  int64_t* seg_first = nullptr;      // max_blocks: records of a piece's first bucket, when split
  int64_t* seg_last = nullptr;       // max_blocks: records of a piece's last bucket, when split
  uint64_t* tail_total = nullptr;    // nbuckets: tail rows of the batch per bucket
  uint64_t* tail_start = nullptr;    // nbuckets: pool position of the bucket's first tail row
  unsigned long long* totals = nullptr;   // [0] records of the new pool, [1] groups
  int* flags = nullptr;              // [0] bucket overflow, [1] key outside its packed range
  int64_t rows = 0;
};

namespace fpdb {

static int32_t acc_op_for(const fpdb_agg_spec& s) {
  bool f = type_is_float(s.input_type);
  bool sg = type_is_signed_int(s.input_type);
  switch (s.kind) {
    case FPDB_AGG_COUNT_ALL: return kAccCountAll;
    case FPDB_AGG_COUNT_VALID: return kAccCountValid;
    case FPDB_AGG_SUM:
    case FPDB_AGG_MEAN: return f ? kAccSumF : kAccSumI;
    case FPDB_AGG_MIN: return f ? kAccMinF : (sg ? kAccMinI : kAccMinU);
    case FPDB_AGG_MAX: return f ? kAccMaxF : (sg ? kAccMaxI : kAccMaxU);
    default: return -1;
  }
}

static fpdb_status gb_alloc(fpdb_group_table* t, uint64_t bytes, void** p) {
  return ctx_alloc(t->ctx, std::max<uint64_t>(bytes, 8), p);
}

}  // namespace fpdb

using namespace fpdb;

extern "C" {

int32_t fpdb_groupby_output_type(fpdb_agg_spec s) {
  switch (s.kind) {
    case FPDB_AGG_COUNT_ALL:
    case FPDB_AGG_COUNT_VALID: return FPDB_TYPE_INT64;
    case FPDB_AGG_SUM:
      if (type_is_float(s.input_type)) return FPDB_TYPE_FLOAT64;
      return type_is_signed_int(s.input_type) ? FPDB_TYPE_INT64 : FPDB_TYPE_UINT64;
    case FPDB_AGG_MIN:
    case FPDB_AGG_MAX: return s.input_type;
    case FPDB_AGG_MEAN: return FPDB_TYPE_FLOAT64;
    default: return 0;
  }
}

}  // extern "C"

namespace fpdb {

static fpdb_status groupby_create_impl(fpdb_context* ctx, const int32_t* key_types, int num_keys,
                                       const int64_t* key_min, const int32_t* key_bits, const fpdb_agg_spec* aggs,
                                       int num_aggs, int64_t max_groups, fpdb_group_table** out) {
  FPDB_CHECK(num_aggs >= 0 && (num_aggs == 0 || aggs), FPDB_ERR_INVALID, "group-by: bad aggregates");
  FPDB_CHECK(num_aggs <= kMaxAggs, FPDB_ERR_INVALID, "group-by: too many aggregates");
  FPDB_CHECK(max_groups >= 1 && max_groups < (int64_t(1) << 32), FPDB_ERR_INVALID,
             "group-by: max_groups must be in [1, 2^32)");
  Context* c = as_ctx(ctx);
  auto* t = new fpdb_group_table();
  t->ctx = c;
  t->num_keys = num_keys;
  t->max_groups = max_groups;
  int shift = 0;
  for (int i = 0; i < num_keys; ++i) {
    t->key_types[i] = key_types[i];
    if (key_bits) {               // packed: caller-given ranges, checked
      t->key_bias[i] = key_min[i];
      t->key_bits[i] = key_bits[i];
      t->key_check = 1;
    } else {                      // one column as is, or the low 32 bits of each of two
      t->key_bias[i] = 0;
      t->key_bits[i] = num_keys == 1 ? 64 : 32;
    }
    t->key_shift[i] = shift;
    shift += t->key_bits[i];
  }
  t->num_aggs = num_aggs;
  std::memset(&t->acc, 0, sizeof(t->acc));
  int nacc = num_aggs;
  for (int a = 0; a < num_aggs; ++a) {
    t->aggs[a] = aggs[a];
    int32_t op = acc_op_for(aggs[a]);
    if (op < 0 || (aggs[a].kind != FPDB_AGG_COUNT_ALL && type_width(aggs[a].input_type) == 0)) {
      delete t;
      return fail(FPDB_ERR_INVALID, "group-by: unsupported aggregate " + std::to_string(a));
    }
    t->acc.op[a] = op;
    t->acc.input[a] = aggs[a].kind == FPDB_AGG_COUNT_ALL ? -1 : a;
    t->hidden_of[a] = -1;
  }
  for (int a = 0; a < num_aggs; ++a) {
    bool needs = aggs[a].kind == FPDB_AGG_MEAN ||
                 (aggs[a].input_nullable && (aggs[a].kind == FPDB_AGG_SUM || aggs[a].kind == FPDB_AGG_MIN ||
                                             aggs[a].kind == FPDB_AGG_MAX));
    if (!needs) continue;
    if (nacc >= kMaxAggs) {
      delete t;
      return fail(FPDB_ERR_INVALID, "group-by: too many aggregates");
    }
    t->acc.op[nacc] = kAccCountValid;
    t->acc.input[nacc] = a;
    t->hidden_of[a] = nacc;
    ++nacc;
  }
  t->acc.n = nacc;
  t->rec_words = 2 + nacc;
  // LDS budget per workgroup (one workgroup of 1024 threads per CU)
  size_t budget = std::min<size_t>(c->max_lds_per_block, 160 * 1024) - 256;   // room for static LDS
  size_t per_slot = 8 * (size_t)(2 + nacc);
  // buckets: a compact workgroup sees about max_groups / nbuckets groups; keep
  // that within a quarter of its slots
  uint32_t cslots = (uint32_t)(budget / per_slot) - 1;
  int64_t nb = kMinBuckets;
  while (nb < kMaxBuckets && nb * (int64_t)(cslots / 4) < max_groups) nb <<= 1;
  t->nbuckets = (int)nb;
  int lg = 0;
  while ((int64_t(1) << lg) < nb) ++lg;
  t->bucket_shift = 64 - lg;
  t->compact_slots = cslots;
  t->compact_lds = (size_t)(cslots + 1) * per_slot;
  // main pass LDS: head index (kHeadIndexSlots x 2 B), per head key its key,
  // batch-relative first row (4 B) and accumulators, per bucket a 64-bit
  // record position; at most a sixth of the index slots hold head keys (the
  // sampler's two-choice placement then rarely drops one)
  // (plus, with at most 5 accumulators, the staged main pass's tail-row stage)
  const size_t stage = nacc >= 1 && nacc <= 5 ? std::max(stage_bytes(t->acc, 1), stage_bytes(t->acc, 2)) : 0;
  const size_t fixed = (size_t)kHeadIndexSlots * 2 + (size_t)nb * 8 + stage;
  const size_t per_head = 8 + 4 + 8 * (size_t)nacc;
  int64_t heads = budget > fixed ? (int64_t)((budget - fixed) / per_head) : 0;
  heads = std::min<int64_t>(heads, kHeadIndexSlots / 6) & ~int64_t(1);   // even: the stage after them stays 16-byte aligned
  t->max_heads = (int)heads;
  t->main_lds = fixed + per_head * (size_t)heads;
  t->count_lds = (size_t)kHeadIndexSlots * 2 + (size_t)heads * 8 + (size_t)nb * 4;
  t->sample_lds = (size_t)kHeadIndexSlots * 4 + (size_t)heads * 12 + (size_t)nb * 4;
  t->max_blocks = std::min(1024, std::max(1, c->num_cus));   // the bucket scan takes up to 1024 workgroups
  if (heads < 64 || cslots < 64 || t->sample_lds > budget) {
    delete t;
    return fail(FPDB_ERR_INVALID, "group-by: too many aggregates for the LDS table");
  }
  fpdb_status s = gb_alloc(t, (uint64_t)nb * 8, reinterpret_cast<void**>(&t->ostart));
  if (s == FPDB_OK) s = gb_alloc(t, (uint64_t)nb * 8, reinterpret_cast<void**>(&t->ocount));
  if (s == FPDB_OK) s = gb_alloc(t, (uint64_t)nb * 8, reinterpret_cast<void**>(&t->nstart));
  if (s == FPDB_OK) s = gb_alloc(t, (uint64_t)nb * 8, reinterpret_cast<void**>(&t->newcount));
  if (s == FPDB_OK) s = gb_alloc(t, (uint64_t)nb * t->max_blocks * 8, reinterpret_cast<void**>(&t->base));
  if (s == FPDB_OK) s = gb_alloc(t, (uint64_t)nb * t->max_blocks * kCountSplit * 4, reinterpret_cast<void**>(&t->hist));
  if (s == FPDB_OK) s = gb_alloc(t, (uint64_t)kHeadIndexSlots * 2, reinterpret_cast<void**>(&t->head_index));
  if (s == FPDB_OK) s = gb_alloc(t, (uint64_t)heads * 8, reinterpret_cast<void**>(&t->head_keys));
  if (s == FPDB_OK) s = gb_alloc(t, (uint64_t)heads * 4, reinterpret_cast<void**>(&t->head_rank));
  if (s == FPDB_OK) s = gb_alloc(t, (uint64_t)nb * 4, reinterpret_cast<void**>(&t->head_count));
  if (s == FPDB_OK) s = gb_alloc(t, 8, reinterpret_cast<void**>(&t->nhead));
  if (s == FPDB_OK)
    s = gb_alloc(t, (uint64_t)t->max_blocks * (1 + nacc) * heads * 8, reinterpret_cast<void**>(&t->hpart));
  if (s == FPDB_OK) s = gb_alloc(t, (uint64_t)nb * 8, reinterpret_cast<void**>(&t->bcount));
  if (s == FPDB_OK) s = gb_alloc(t, (uint64_t)t->max_blocks * 8, reinterpret_cast<void**>(&t->seg_first));
  if (s == FPDB_OK) s = gb_alloc(t, (uint64_t)t->max_blocks * 8, reinterpret_cast<void**>(&t->seg_last));
  if (s == FPDB_OK) s = gb_alloc(t, (uint64_t)nb * 8, reinterpret_cast<void**>(&t->tail_total));
  if (s == FPDB_OK) s = gb_alloc(t, (uint64_t)nb * 8, reinterpret_cast<void**>(&t->tail_start));
  if (s == FPDB_OK) s = gb_alloc(t, 32, reinterpret_cast<void**>(&t->totals));
  if (s == FPDB_OK) s = gb_alloc(t, 8, reinterpret_cast<void**>(&t->flags));
  if (s == FPDB_OK) s = gb_alloc(t, (uint64_t)(kSampleTable + 1) * 8, reinterpret_cast<void**>(&t->sample_keys));
  if (s == FPDB_OK) s = gb_alloc(t, (uint64_t)(kSampleTable + 1) * 4, reinterpret_cast<void**>(&t->sample_cnt));
  if (s == FPDB_OK) s = launch_fill(c, t->sample_keys, 0xFF, (uint64_t)(kSampleTable + 1) * 8);
  if (s == FPDB_OK) s = launch_fill(c, t->sample_cnt, 0, (uint64_t)(kSampleTable + 1) * 4);
  if (s == FPDB_OK) s = gb_alloc(t, (uint64_t)kSampleListMax * 8, reinterpret_cast<void**>(&t->list_keys));
  if (s == FPDB_OK) s = gb_alloc(t, (uint64_t)kSampleListMax * 4, reinterpret_cast<void**>(&t->list_cnt));
  if (s == FPDB_OK) s = gb_alloc(t, (uint64_t)(kCountBins + 1) * 4, reinterpret_cast<void**>(&t->list_hist));
  if (s == FPDB_OK) s = launch_fill(c, t->list_hist, 0, (uint64_t)(kCountBins + 1) * 4);
  if (s == FPDB_OK) s = launch_fill(c, t->ostart, 0, (uint64_t)nb * 8);
  if (s == FPDB_OK) s = launch_fill(c, t->ocount, 0, (uint64_t)nb * 8);
  if (s == FPDB_OK) s = launch_fill(c, t->totals, 0, 32);
  if (s == FPDB_OK) s = launch_fill(c, t->flags, 0, 8);
  if (s != FPDB_OK) {
    fpdb_groupby_destroy(t);
    return s;
  }
  *out = t;
  return FPDB_OK;
}

template <int NACC, int KW, bool V8>
static void launch_main_kv(Context* c, fpdb_group_table* t, int blocks, const GKeySpec& k, const ColSet& vals,
                           const GbGeom& G) {
  if constexpr (NACC > 0)
    hipLaunchKernelGGL((gb_main_staged_kernel<NACC, KW, V8>), dim3(blocks), dim3(kGroupThreads), t->main_lds,
                       c->stream, k, vals, t->acc, G, t->head_index, t->head_keys, t->nhead, t->base, t->tail_start,
                       t->tpool, t->hpart);
  else
    hipLaunchKernelGGL((gb_main_kernel<NACC, KW, V8>), dim3(blocks), dim3(kGroupThreads), t->main_lds, c->stream, k,
                       vals, t->acc, G, t->head_index, t->head_keys, t->nhead, t->base, t->tail_start, t->spare,
                       t->hpart);
}

// The key width of the direct key path (one unpacked column of 4 or 8 bytes),
// or 0 for the generic key walk.
static int gb_key_width(const fpdb_group_table* t, const GKeySpec& k) {
  if (k.n != 1 || k.check || k.bits[0] != 64 || k.shift[0] != 0 || k.bias[0] != 0) return 0;
  return (k.w[0] == 4 || k.w[0] == 8) ? k.w[0] : 0;
}

// True when every aggregate input is an 8-byte column without nulls.
static bool gb_values_8(const fpdb_group_table* t, const ColSet& vals) {
  for (int a = 0; a < t->acc.n; ++a) {
    const int32_t op = t->acc.op[a];
    if (op == kAccCountAll) continue;
    const int in = t->acc.input[a];
    if (vals.width[in] != 8 || vals.validity[in]) return false;
    if (op != kAccCountValid && vals.type[in] != FPDB_TYPE_INT64 && vals.type[in] != FPDB_TYPE_UINT64 &&
        vals.type[in] != FPDB_TYPE_FLOAT64 && vals.type[in] != FPDB_TYPE_DATE64)
      return false;
  }
  return true;
}

template <int NACC>
static void launch_main(Context* c, fpdb_group_table* t, int blocks, const GKeySpec& k, const ColSet& vals,
                        const GbGeom& G) {
  const int kw = gb_key_width(t, k);
  if (NACC > 0 && kw == 4 && gb_values_8(t, vals))
    launch_main_kv<NACC, 4, true>(c, t, blocks, k, vals, G);
  else if (NACC > 0 && kw == 8 && gb_values_8(t, vals))
    launch_main_kv<NACC, 8, true>(c, t, blocks, k, vals, G);
  else
    launch_main_kv<NACC, 0, false>(c, t, blocks, k, vals, G);
}
template <int NACC>
static void launch_compact(Context* c, fpdb_group_table* t, int rk, int rr, int64_t row_base) {
  if constexpr (NACC > 0) {
    // raw tail rows: balanced pieces, then the merge of buckets split over pieces
    const int pieces = t->max_blocks;
    hipLaunchKernelGGL(gb_compact_piece_kernel<NACC>, dim3(pieces), dim3(kGroupThreads), t->compact_lds, c->stream,
                       t->acc, t->compact_slots, t->rec_words, t->nbuckets, t->totals, t->pool, t->ostart, t->ocount,
                       t->spare, t->nstart, t->newcount, t->tpool, t->tail_start, t->tail_total, rk, rr, row_base,
                       t->bcount, t->seg_first, t->seg_last, t->flags);
    hipLaunchKernelGGL(gb_compact_merge_kernel<NACC>, dim3(t->nbuckets), dim3(kGroupThreads), t->compact_lds,
                       c->stream, t->acc, t->compact_slots, t->rec_words, pieces, t->totals, t->spare, t->nstart,
                       t->ostart, t->ocount, t->newcount, t->tail_total, t->bcount, t->seg_first, t->seg_last,
                       t->flags);
  } else {
    hipLaunchKernelGGL(gb_compact_kernel<NACC>, dim3(t->nbuckets), dim3(kGroupThreads), t->compact_lds, c->stream,
                       t->acc, t->compact_slots, t->rec_words, t->pool, t->ostart, t->ocount, t->spare, t->nstart,
                       t->newcount, nullptr, t->tail_start, t->tail_total, rk, rr, row_base, t->flags);
  }
}

#define FPDB_GB_DISPATCH(FN, ...)                 \
  switch (t->acc.n) {                             \
    case 1: FN<1>(__VA_ARGS__); break;            \
    case 2: FN<2>(__VA_ARGS__); break;            \
    case 3: FN<3>(__VA_ARGS__); break;            \
    case 4: FN<4>(__VA_ARGS__); break;            \
    case 5: FN<5>(__VA_ARGS__); break;            \
    default: FN<0>(__VA_ARGS__); break;           \
  }

// One batch of n <= kMaxBatchRows rows: sample, count, offsets, main, merge,
// compact, all on the context stream (no host synchronization unless the
// worst-case size of the next pool does not fit the HBM budget).
static fpdb_status groupby_consume_batch(Context* c, fpdb_group_table* t, const GKeySpec& k, const ColSet& vals,
                                         int64_t n) {
  const int blocks = t->max_blocks;
  GbGeom G;
  G.n = n;
  G.row_base = t->rows;
  G.max_heads = t->max_heads;
  G.nbuckets = t->nbuckets;
  G.bucket_shift = t->bucket_shift;
  G.rec_words = t->rec_words;
  // The next pool holds at most: the groups so far (at most the records of
  // the current pool), every head key and every row of the batch.
  const int64_t bound = t->pool_records_bound + t->max_heads + n;
  bool exact = false;
  // tail rows raw in the tail pool (the staged main pass: 1 to 5 accumulators)
  const bool separate = t->acc.n >= 1 && t->acc.n <= 5;
  const int kw = gb_key_width(t, k);
  int nv = 0;
  for (int a = 0; a < t->acc.n; ++a) nv += t->acc.op[a] == kAccCountAll ? 0 : 1;
  const int rr = (kw == 4 ? 1 : 2) + nv;
  const int rk = kw == 4 ? (k.sgn[0] ? 3 : 1) : 2;
  bool exact_tail = false;
  if (separate && n * rr > t->tpool_words) {
    if (t->tpool) FPDB_TRY(ctx_free(c, t->tpool));
    t->tpool = nullptr;
    t->tpool_words = 0;
    fpdb_status s = gb_alloc(t, (uint64_t)n * rr * 8, reinterpret_cast<void**>(&t->tpool));
    if (s == FPDB_ERR_BUDGET) {
      (void)fpdb_last_error();
      exact_tail = true;   // size it from the count pass instead
    } else {
      FPDB_TRY(s);
      t->tpool_words = n * rr;
    }
  }
  if (bound > t->spare_cap) {
    if (t->spare) FPDB_TRY(ctx_free(c, t->spare));
    t->spare = nullptr;
    t->spare_cap = 0;
    const int64_t want = std::max<int64_t>(bound, 1024);
    fpdb_status s = gb_alloc(t, (uint64_t)want * t->rec_words * 8, reinterpret_cast<void**>(&t->spare));
    if (s == FPDB_ERR_BUDGET) {
      (void)fpdb_last_error();
      exact = true;   // size it from the count pass instead
    } else {
      FPDB_TRY(s);
      t->spare_cap = want;
    }
  }
  {
    {
      ProfScope ps(c, "groupby_sample");
      hipLaunchKernelGGL(gb_sample_count_kernel, dim3(kSampleBlocks), dim3(kSampleBlockRows), 0, c->stream, k, G,
                         t->sample_keys, t->sample_cnt);
      hipLaunchKernelGGL(gb_sample_gather_kernel, dim3(kSampleTable / 1024 + 1), dim3(1024), 0, c->stream,
                         t->sample_keys, t->sample_cnt, t->list_keys, t->list_cnt, t->list_hist);
      hipLaunchKernelGGL(gb_sample_select_kernel, dim3(1), dim3(1024), t->sample_lds, c->stream, G, t->list_keys,
                         t->list_cnt, t->list_hist, t->head_index, t->head_keys, t->head_rank, t->head_count,
                         t->nhead);
    }
    ProfScope ps(c, "groupby_count");
    auto count_kernel = kw == 8 ? gb_count_kernel<8> : (kw == 4 ? gb_count_kernel<4> : gb_count_kernel<0>);
    hipLaunchKernelGGL(count_kernel, dim3(blocks * kCountSplit), dim3(kGroupThreads), t->count_lds, c->stream, k, G,
                       t->head_index, t->head_keys, t->nhead, t->hist, t->flags + 1);
    hipLaunchKernelGGL(gb_bucket_scan_kernel, dim3((unsigned)ceil_div(t->nbuckets, 16)), dim3(1024), 0, c->stream,
                       t->hist, blocks, t->nbuckets, t->base, t->tail_total);
    hipLaunchKernelGGL(gb_bucket_offsets_kernel, dim3(1), dim3(1024), 0, c->stream, t->nbuckets, separate ? 1 : 0,
                       t->ocount,
                       t->head_count, t->tail_total, t->nstart, t->newcount, t->tail_start, t->totals);
    FPDB_HIP(hipGetLastError());
  }
  if (exact || exact_tail) {
    uint64_t tot[3] = {0, 0, 0};
    FPDB_TRY(mail_read(c, t->totals, 3, tot));
    if (exact) {
      const int64_t want = std::max<int64_t>((int64_t)tot[0], 1024);
      FPDB_TRY(gb_alloc(t, (uint64_t)want * t->rec_words * 8, reinterpret_cast<void**>(&t->spare)));
      t->spare_cap = want;
    }
    if (exact_tail) {
      const int64_t want = std::max<int64_t>((int64_t)tot[2], 1) * rr;
      FPDB_TRY(gb_alloc(t, (uint64_t)want * 8, reinterpret_cast<void**>(&t->tpool)));
      t->tpool_words = want;
    }
  }
  {
    ProfScope ps(c, "groupby_main");
    FPDB_GB_DISPATCH(launch_main, c, t, blocks, k, vals, G);
    FPDB_HIP(hipGetLastError());
  }
  {
    ProfScope ps(c, "groupby_compact");
    hipLaunchKernelGGL(gb_head_merge_kernel, dim3((unsigned)ceil_div(t->max_heads, 64), (unsigned)(1 + t->acc.n)),
                       dim3(64 * kMergeWaves), 0, c->stream, t->acc, G, blocks, t->head_keys, t->head_rank,
                       t->nhead, t->hpart, t->nstart, t->ocount, t->spare);
    FPDB_GB_DISPATCH(launch_compact, c, t, rk, rr, G.row_base);
    FPDB_HIP(hipGetLastError());
  }
  std::swap(t->pool, t->spare);
  std::swap(t->pool_cap, t->spare_cap);
  // groups after the batch: at most the compact workgroups' slots, and at
  // most the old groups plus the batch's rows
  t->pool_records_bound = std::min<int64_t>(t->pool_records_bound + n, (int64_t)t->nbuckets * (t->compact_slots + 1));
  t->rows += n;
  return FPDB_OK;
}

}  // namespace fpdb

extern "C" {

fpdb_status fpdb_groupby_create(fpdb_context* ctx, const int32_t* key_types, int num_keys,
                                const fpdb_agg_spec* aggs, int num_aggs, int64_t max_groups,
                                fpdb_group_table** out) {
  FPDB_CHECK(ctx && out, FPDB_ERR_INVALID, "NULL argument");
  *out = nullptr;
  FPDB_CHECK(num_keys >= 0 && num_keys <= 2, FPDB_ERR_INVALID,
             "group-by: zero, one or two key columns (more through fpdb_groupby_create_packed)");
  FPDB_CHECK(key_types || num_keys == 0, FPDB_ERR_INVALID, "group-by: NULL key types");
  for (int i = 0; i < num_keys; ++i) {
    int w = type_width(key_types[i]);
    FPDB_CHECK(w > 0 && !type_is_float(key_types[i]), FPDB_ERR_INVALID, "group-by: keys must be integers");
    FPDB_CHECK(num_keys == 1 || w <= 4, FPDB_ERR_INVALID,
               "group-by: two-column keys need columns of at most 4 bytes");
  }
  return groupby_create_impl(ctx, key_types, num_keys, nullptr, nullptr, aggs, num_aggs, max_groups, out);
}

fpdb_status fpdb_groupby_create_packed(fpdb_context* ctx, const int32_t* key_types, int num_keys,
                                       const int64_t* key_min, const int32_t* key_bits, const fpdb_agg_spec* aggs,
                                       int num_aggs, int64_t max_groups, fpdb_group_table** out) {
  FPDB_CHECK(ctx && out, FPDB_ERR_INVALID, "NULL argument");
  *out = nullptr;
  FPDB_CHECK(num_keys >= 1 && num_keys <= kMaxGroupKeys, FPDB_ERR_INVALID, "group-by: one to four packed key columns");
  FPDB_CHECK(key_types && key_min && key_bits, FPDB_ERR_INVALID, "group-by: NULL key description");
  int total = 0;
  for (int i = 0; i < num_keys; ++i) {
    int w = type_width(key_types[i]);
    FPDB_CHECK(w > 0 && !type_is_float(key_types[i]), FPDB_ERR_INVALID, "group-by: keys must be integers");
    FPDB_CHECK(key_bits[i] >= 1 && key_bits[i] <= 64, FPDB_ERR_INVALID, "group-by: key bits must be in [1, 64]");
    total += key_bits[i];
  }
  FPDB_CHECK(total <= 64, FPDB_ERR_INVALID, "group-by: packed key wider than 64 bits");
  return groupby_create_impl(ctx, key_types, num_keys, key_min, key_bits, aggs, num_aggs, max_groups, out);
}

fpdb_status fpdb_column_min_max(fpdb_context* ctx, const fpdb_column* column, int64_t num_rows, int64_t* min,
                                int64_t* max, int64_t* num_valid) {
  FPDB_CHECK(ctx && column && min && max, FPDB_ERR_INVALID, "NULL argument");
  FPDB_CHECK(!type_is_float(column->type) && type_width(column->type) > 0, FPDB_ERR_INVALID,
             "min/max: integer column needed");
  Context* c = as_ctx(ctx);
  ColSet cs;
  FPDB_TRY(make_colset(column, 1, num_rows, &cs, "min/max"));
  unsigned long long* d = nullptr;
  FPDB_TRY(ctx_alloc(c, 24, reinterpret_cast<void**>(&d)));
  unsigned long long h[3] = {~0ULL, 0ULL, 0ULL};
  hipError_t e = hipSuccess;
  if (num_rows > 0) {
    const int grid = std::min(grid_for(c, num_rows, 256, 4), kMaxPartials / 3);
    hipLaunchKernelGGL(range_kernel, dim3(grid), dim3(256), 0, c->stream, cs, num_rows, c->d_partials);
    hipLaunchKernelGGL(range_fold_kernel, dim3(1), dim3(1024), 0, c->stream, c->d_partials, grid, d);
    e = hipGetLastError();
  }
  if (e == hipSuccess && num_rows > 0) e = hipMemcpyAsync(h, d, 24, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  ctx_free(c, d);
  FPDB_CHECK(e == hipSuccess, FPDB_ERR_HIP, hipGetErrorString(e));
  const bool sgn = type_is_signed_int(column->type);
  if (num_valid) *num_valid = (int64_t)h[2];
  if (h[2] == 0) {
    *min = 0;
    *max = 0;
  } else {
    *min = (int64_t)(sgn ? (h[0] ^ (1ULL << 63)) : h[0]);
    *max = (int64_t)(sgn ? (h[1] ^ (1ULL << 63)) : h[1]);
  }
  return FPDB_OK;
}

fpdb_status fpdb_groupby_destroy(fpdb_group_table* t) {
  if (!t) return FPDB_OK;
  void* bufs[] = {t->pool, t->spare, t->tpool, t->sample_keys, t->sample_cnt, t->list_keys, t->list_cnt,
                  t->list_hist, t->ostart, t->ocount, t->nstart, t->newcount, t->base, t->hist,
                  t->head_index, t->head_keys, t->head_rank, t->head_count, t->nhead, t->hpart,
                  t->bcount, t->seg_first, t->seg_last, t->tail_total, t->tail_start, t->totals, t->flags};
  for (void* b : bufs)
    if (b) ctx_free(t->ctx, b);
  delete t;
  return FPDB_OK;
}

fpdb_status fpdb_groupby_consume(fpdb_context* ctx, fpdb_group_table* t, const fpdb_column* keys,
                                 const fpdb_column* values, int64_t n) {
  FPDB_CHECK(ctx && t, FPDB_ERR_INVALID, "NULL argument");
  FPDB_CHECK(n >= 0, FPDB_ERR_INVALID, "negative row count");
  Context* c = as_ctx(ctx);
  GKeySpec k;
  std::memset(&k, 0, sizeof(k));
  k.n = t->num_keys;
  k.check = t->key_check;
  for (int i = 0; i < t->num_keys; ++i) {
    FPDB_CHECK(keys && keys[i].length >= n && (n == 0 || keys[i].values), FPDB_ERR_INVALID,
               "group-by: key column too short");
    FPDB_CHECK(type_width(keys[i].type) == type_width(t->key_types[i]) &&
                   type_is_signed_int(keys[i].type) == type_is_signed_int(t->key_types[i]),
               FPDB_ERR_INVALID, "group-by: key column type differs from the table's");
    FPDB_CHECK(keys[i].validity == nullptr, FPDB_ERR_INVALID, "group-by: null group keys are not supported");
    k.v[i] = static_cast<const uint8_t*>(keys[i].values);
    k.off[i] = keys[i].offset;
    k.w[i] = type_width(keys[i].type);
    k.sgn[i] = type_is_signed_int(keys[i].type) ? 1 : 0;
    k.bias[i] = t->key_bias[i];
    k.bits[i] = t->key_bits[i];
    k.shift[i] = t->key_shift[i];
  }
  ColSet vals;
  std::memset(&vals, 0, sizeof(vals));
  vals.n = t->num_aggs;
  for (int a = 0; a < t->num_aggs; ++a) {
    if (t->aggs[a].kind == FPDB_AGG_COUNT_ALL) continue;
    FPDB_CHECK(values && values[a].length >= n && (n == 0 || values[a].values), FPDB_ERR_INVALID,
               "group-by: value column " + std::to_string(a) + " too short");
    FPDB_CHECK(values[a].type == t->aggs[a].input_type, FPDB_ERR_INVALID,
               "group-by: value column " + std::to_string(a) + " type differs from the aggregate's");
    FPDB_CHECK(values[a].validity == nullptr || t->aggs[a].input_nullable, FPDB_ERR_INVALID,
               "group-by: value column " + std::to_string(a) + " has nulls but the aggregate was created non-nullable");
    vals.values[a] = static_cast<const uint8_t*>(values[a].values);
    vals.validity[a] = values[a].validity;
    vals.offset[a] = values[a].offset;
    vals.width[a] = type_width(values[a].type);
    vals.type[a] = values[a].type;
  }
  if (n == 0) return FPDB_OK;
  FPDB_CHECK(t->rows + n < (int64_t(1) << 62), FPDB_ERR_INVALID, "group-by: too many rows");
  // batches beyond kMaxBatchRows rows are consumed in pieces (rows are
  // batch-relative 32-bit values in the main pass's LDS)
  for (int64_t done = 0; done < n; done += kMaxBatchRows) {
    const int64_t m = std::min<int64_t>(kMaxBatchRows, n - done);
    GKeySpec kp = k;
    ColSet vp = vals;
    for (int i = 0; i < kp.n; ++i) kp.off[i] += done;
    for (int a = 0; a < vp.n; ++a) vp.offset[a] += done;
    FPDB_TRY(groupby_consume_batch(c, t, kp, vp, m));
  }
  return FPDB_OK;
}

fpdb_status fpdb_groupby_num_groups(fpdb_context* ctx, fpdb_group_table* t, int64_t* num_groups) {
  FPDB_CHECK(ctx && t && num_groups, FPDB_ERR_INVALID, "NULL argument");
  Context* c = as_ctx(ctx);
  unsigned long long groups[2] = {0, 0};
  int flags[2] = {0, 0};
  hipLaunchKernelGGL(gb_sum_groups_kernel, dim3(1), dim3(1024), 0, c->stream, t->ocount, t->nbuckets, t->totals + 1);
  FPDB_HIP(hipGetLastError());
  FPDB_HIP(hipMemcpyAsync(groups, t->totals, 16, hipMemcpyDeviceToHost, c->stream));
  FPDB_HIP(hipMemcpyAsync(flags, t->flags, 8, hipMemcpyDeviceToHost, c->stream));
  FPDB_HIP(hipStreamSynchronize(c->stream));
  FPDB_CHECK(flags[1] == 0, FPDB_ERR_INVALID,
             "group-by: a key value lies outside the range the packed key was created for");
  FPDB_CHECK(flags[0] == 0 && (int64_t)groups[1] <= t->max_groups, FPDB_ERR_CAPACITY,
             "group-by: more groups than max_groups");
  *num_groups = (int64_t)groups[1];
  return FPDB_OK;
}

fpdb_status fpdb_groupby_finalize(fpdb_context* ctx, fpdb_group_table* t, fpdb_column* out_keys,
                                  fpdb_column* out_aggs, int64_t* num_groups) {
  FPDB_CHECK(ctx && t && num_groups, FPDB_ERR_INVALID, "NULL argument");
  Context* c = as_ctx(ctx);
  int64_t ng = 0;
  FPDB_TRY(fpdb_groupby_num_groups(ctx, t, &ng));
  *num_groups = ng;
  GOutSpec o;
  std::memset(&o, 0, sizeof(o));
  o.num_keys = t->num_keys;
  o.num_aggs = t->num_aggs;
  for (int i = 0; i < t->num_keys; ++i) {
    FPDB_CHECK(out_keys && type_width(out_keys[i].type) == type_width(t->key_types[i]), FPDB_ERR_INVALID,
               "group-by: output key column type differs");
    FPDB_CHECK(ng == 0 || out_keys[i].values, FPDB_ERR_INVALID, "group-by: NULL output key buffer");
    o.key_values[i] = static_cast<uint8_t*>(out_keys[i].values);
    o.key_w[i] = type_width(out_keys[i].type);
    o.key_off[i] = out_keys[i].offset;
    o.key_bias[i] = t->key_bias[i];
    o.key_bits[i] = t->key_bits[i];
    o.key_shift[i] = t->key_shift[i];
  }
  for (int a = 0; a < t->num_aggs; ++a) {
    int32_t ot = fpdb_groupby_output_type(t->aggs[a]);
    FPDB_CHECK(out_aggs && type_width(out_aggs[a].type) == type_width(ot), FPDB_ERR_INVALID,
               "group-by: output aggregate column " + std::to_string(a) + " has the wrong width");
    FPDB_CHECK(ng == 0 || out_aggs[a].values, FPDB_ERR_INVALID, "group-by: NULL output aggregate buffer");
    FPDB_CHECK(t->hidden_of[a] < 0 || out_aggs[a].validity, FPDB_ERR_INVALID,
               "group-by: nullable aggregate needs an output validity bitmap");
    o.agg_values[a] = static_cast<uint8_t*>(out_aggs[a].values);
    o.agg_validity[a] = out_aggs[a].validity;
    o.agg_off[a] = out_aggs[a].offset;
    o.agg_w[a] = type_width(ot);
    o.agg_count_acc[a] = t->hidden_of[a];
    o.agg_is_f32[a] = (t->aggs[a].input_type == FPDB_TYPE_FLOAT32 &&
                       (t->aggs[a].kind == FPDB_AGG_MIN || t->aggs[a].kind == FPDB_AGG_MAX)) ? 1 : 0;
    o.agg_mean[a] = t->aggs[a].kind != FPDB_AGG_MEAN ? 0 : (type_is_float(t->aggs[a].input_type) ? 2 : 1);
    if (out_aggs[a].validity) {
      FPDB_CHECK((out_aggs[a].offset & 31) == 0, FPDB_ERR_INVALID, "output validity offset must be a multiple of 32");
      FPDB_TRY(launch_fill(c, out_aggs[a].validity + out_aggs[a].offset / 8, 0, (uint64_t)((ng + 31) / 32) * 4));
    }
  }
  if (ng == 0) return FPDB_OK;
  int64_t words = ceil_div(t->rows, 64);
  int64_t chunks = ceil_div(words, kScanChunk);
  unsigned long long* bits = nullptr;
  uint32_t* chunk_sums = nullptr;
  uint32_t* word_prefix = nullptr;
  fpdb_status s = ctx_alloc(c, (uint64_t)words * 8, reinterpret_cast<void**>(&bits));
  if (s == FPDB_OK) s = ctx_alloc(c, (uint64_t)chunks * 4, reinterpret_cast<void**>(&chunk_sums));
  if (s == FPDB_OK) s = ctx_alloc(c, (uint64_t)words * 4, reinterpret_cast<void**>(&word_prefix));
  if (s == FPDB_OK) s = launch_fill(c, bits, 0, (uint64_t)words * 8);
  if (s == FPDB_OK) {
    hipLaunchKernelGGL(gb_mark_kernel, dim3(t->nbuckets), dim3(256), 0, c->stream, t->pool, t->ostart, t->ocount,
                       t->rec_words, bits);
    hipLaunchKernelGGL(popcount_chunk_kernel, dim3((unsigned)chunks), dim3(1024), 0, c->stream, bits, words,
                       chunk_sums);
    hipLaunchKernelGGL(chunk_scan_kernel, dim3(1), dim3(1024), 0, c->stream, chunk_sums, chunks);
    hipLaunchKernelGGL(word_prefix_kernel, dim3((unsigned)chunks), dim3(1024), 0, c->stream, bits, words,
                       chunk_sums, word_prefix);
    hipLaunchKernelGGL(gb_emit_kernel, dim3(t->nbuckets), dim3(256), 0, c->stream, t->pool, t->ostart, t->ocount,
                       t->rec_words, bits, word_prefix, o);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) s = fail(FPDB_ERR_HIP, hipGetErrorString(e));
    if (s == FPDB_OK) {
      e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) s = fail(FPDB_ERR_HIP, hipGetErrorString(e));
    }
  }
  if (bits) ctx_free(c, bits);
  if (chunk_sums) ctx_free(c, chunk_sums);
  if (word_prefix) ctx_free(c, word_prefix);
  return s;
}

}  // extern "C"

// ---- averages from intermediate sums and counts ------------------------------

namespace fpdb {

// One wave per 64 rows: quotient, validity word from a ballot.
__global__ __launch_bounds__(256) void divide_kernel(ColSet in, int64_t n, double* __restrict__ out,
                                                     uint8_t* __restrict__ out_validity, int64_t out_off,
                                                     int* div_zero) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const bool int_num = !(in.type[0] == FPDB_TYPE_FLOAT32 || in.type[0] == FPDB_TYPE_FLOAT64);
  const bool int_den = !(in.type[1] == FPDB_TYPE_FLOAT32 || in.type[1] == FPDB_TYPE_FLOAT64);
  bool zero = false;
  for (int64_t c = wave; c < (n + 63) / 64; c += nwaves) {
    const int64_t r = c * 64 + lane;
    bool valid = false;
    double q = 0.0;
    if (r < n) {
      const int64_t i0 = r + in.offset[0], i1 = r + in.offset[1];
      valid = valid_bit(in.validity[0], i0) && valid_bit(in.validity[1], i1);
      auto as_f64 = [&](int k, int64_t i) -> double {
        if (in.type[k] == FPDB_TYPE_FLOAT64) return reinterpret_cast<const double*>(in.values[k])[i];
        if (in.type[k] == FPDB_TYPE_FLOAT32) return (double)reinterpret_cast<const float*>(in.values[k])[i];
        if (type_is_signed_int_d(in.type[k])) return (double)load_sext(in.values[k], in.width[k], i);
        return (double)load_zext(in.values[k], in.width[k], i);
      };
      if (valid) {
        if (int_num && int_den) {
          const bool sn = type_is_signed_int_d(in.type[0]), sd = type_is_signed_int_d(in.type[1]);
          const int64_t a = sn ? load_sext(in.values[0], in.width[0], i0) : (int64_t)load_zext(in.values[0], in.width[0], i0);
          const int64_t b = sd ? load_sext(in.values[1], in.width[1], i1) : (int64_t)load_zext(in.values[1], in.width[1], i1);
          if (b == 0) {
            zero = true;
          } else {
            q = (double)(b == -1 ? (int64_t)(0 - (uint64_t)a) : a / b);
          }
        } else {
          q = as_f64(0, i0) / as_f64(1, i1);
        }
      }
      out[r] = q;
    }
    const uint64_t word = __ballot(valid);
    if (out_validity && lane < 2) {
      const int64_t bit0 = c * 64 + out_off + 32 * lane;   // out_off is a multiple of 32
      if (c * 64 + 32 * lane < n)
        reinterpret_cast<uint32_t*>(out_validity)[bit0 >> 5] = (uint32_t)(word >> (32 * lane));
    }
  }
  if (zero) *div_zero = 1;
}

}  // namespace fpdb

extern "C" {

fpdb_status fpdb_divide_to_float64(fpdb_context* ctx, const fpdb_column* num, const fpdb_column* den,
                                   int64_t num_rows, fpdb_column* out) {
  FPDB_CHECK(ctx && num && den && out, FPDB_ERR_INVALID, "NULL argument");
  FPDB_CHECK(out->type == FPDB_TYPE_FLOAT64 && (num_rows == 0 || out->values), FPDB_ERR_INVALID,
             "divide: the output must be a float64 column");
  FPDB_CHECK(out->length >= num_rows, FPDB_ERR_INVALID, "divide: output column too short");
  FPDB_CHECK(out->validity == nullptr || (out->offset & 31) == 0, FPDB_ERR_INVALID,
             "divide: output validity offset must be a multiple of 32");
  fpdb_column both[2] = {*num, *den};
  ColSet in;
  FPDB_TRY(make_colset(both, 2, num_rows, &in, "divide"));
  FPDB_CHECK(out->validity || (!num->validity && !den->validity), FPDB_ERR_INVALID,
             "divide: nullable inputs need an output validity bitmap");
  if (num_rows == 0) return FPDB_OK;
  Context* c = as_ctx(ctx);
  int* flag = reinterpret_cast<int*>(c->d_counter + 2);
  FPDB_TRY(launch_fill(c, flag, 0, 4));
  const int grid = grid_for(c, num_rows, 256, 8);
  hipLaunchKernelGGL(divide_kernel, dim3(grid), dim3(256), 0, c->stream, in, num_rows,
                     static_cast<double*>(out->values) + out->offset, out->validity, out->offset, flag);
  FPDB_HIP(hipGetLastError());
  int zero = 0;
  FPDB_HIP(hipMemcpyAsync(&zero, flag, 4, hipMemcpyDeviceToHost, c->stream));
  FPDB_HIP(hipStreamSynchronize(c->stream));
  FPDB_CHECK(!zero, FPDB_ERR_INVALID, "divide by zero");
  return FPDB_OK;
}

}  // extern "C"
