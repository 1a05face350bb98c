// Hash group-by aggregation on gfx950.
//
// Consume: every workgroup owns a contiguous range of rows and a hash table in
// LDS (keys, one 8-byte accumulator per aggregate, first row).  Rows whose key
// finds or claims an LDS slot within a few probes are aggregated with LDS
// atomics, which absorbs the heavy head of skewed (Zipf) keys; the remaining
// rows go straight to the global open-addressing table in HBM.  At the end the
// workgroup merges its LDS table into the global table.
//
// Integer aggregates are exact (wrapping int64 arithmetic is associative), so
// they are independent of the order in which rows arrive.  FP64 sums are
// accumulated with hardware FP64 atomics.
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
constexpr int kLdsProbes = 8;
constexpr int kGroupThreads = 1024;

__device__ __forceinline__ bool type_is_signed_int_d(int32_t t) {
  return t == FPDB_TYPE_INT8 || t == FPDB_TYPE_INT16 || t == FPDB_TYPE_INT32 || t == FPDB_TYPE_INT64 ||
         t == FPDB_TYPE_DATE32 || t == FPDB_TYPE_DATE64;
}

__device__ __forceinline__ int64_t ceil_div_d(int64_t a, int64_t b) { return (a + b - 1) / b; }

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

// Applies one operand; works on LDS and global addresses alike (the compiler
// picks ds_* or global_* atomics from the address space).
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

// The global table has one replica per XCD: a workgroup aggregates into the
// replica of the XCD it runs on, so the cache lines its atomics touch stay in
// that XCD's L2 instead of bouncing between L2s.  Replicas 1.. are merged into
// replica 0 before groups are counted or emitted.  Any workgroup may use any
// replica (all atomics are device-coherent); the XCD choice is only locality.
constexpr int kReplicas = 8;

struct GlobalTable {
  uint64_t* keys;        // nrep x cap
  uint64_t* acc;         // nrep x nacc x (cap + 1); slot cap holds the key equal to kEmptyKey
  int64_t* first;        // nrep x (cap + 1)
  uint64_t mask;         // cap - 1
  int64_t stride;        // cap + 1
  int32_t nacc;
  int32_t nrep;
  unsigned long long* groups;  // number of claimed slots of replica 0
  int* overflow;
  int* key_range;              // set when a packed key value lies outside its range
};

__device__ __forceinline__ uint64_t* gkey(const GlobalTable& g, int rep, int64_t s) {
  return g.keys + (size_t)rep * (g.mask + 1) + s;
}
__device__ __forceinline__ uint64_t* gacc(const GlobalTable& g, int rep, int a, int64_t s) {
  return g.acc + ((size_t)rep * g.nacc + a) * g.stride + s;
}
__device__ __forceinline__ int64_t* gfirst(const GlobalTable& g, int rep, int64_t s) {
  return g.first + (size_t)rep * g.stride + s;
}

__device__ __forceinline__ int xcc_id() {
  unsigned v;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
  return (int)(v & 0xF);
}

// Finds or claims the slot of key k in a replica (cap for the reserved key value).
__device__ __forceinline__ int64_t global_slot(const GlobalTable& g, int rep, uint64_t k) {
  if (k == kEmptyKey) return (int64_t)g.mask + 1;
  uint64_t s = fmix64(k) & g.mask;
  for (uint64_t p = 0; p <= g.mask; ++p) {
    uint64_t* kp = gkey(g, rep, s);
    uint64_t cur = __hip_atomic_load(kp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == k) return (int64_t)s;
    if (cur == kEmptyKey) {
      unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long*>(kp), kEmptyKey, k);
      if (prev == kEmptyKey) {
        if (rep == 0) atomicAdd(g.groups, 1ULL);
        return (int64_t)s;
      }
      if (prev == k) return (int64_t)s;
    }
    s = (s + 1) & g.mask;
  }
  *g.overflow = 1;
  return -1;
}

template <int NACC>
__global__ __launch_bounds__(kGroupThreads) void groupby_consume_kernel(GKeySpec keys, ColSet vals, AccSpec A,
                                                                        int64_t n, int64_t row_base,
                                                                        GlobalTable g, int lds_slots) {
  extern __shared__ uint64_t lds[];
  const int nacc = NACC > 0 ? NACC : A.n;
  uint64_t* lkeys = lds;
  uint64_t* lacc = lds + lds_slots;                       // nacc x lds_slots
  int64_t* lfirst = reinterpret_cast<int64_t*>(lds + (size_t)lds_slots * (1 + nacc));
  const uint32_t lmask = (uint32_t)lds_slots - 1;
  for (int s = threadIdx.x; s < lds_slots; s += blockDim.x) {
    lkeys[s] = kEmptyKey;
    lfirst[s] = LLONG_MAX;
    for (int a = 0; a < nacc; ++a) lacc[(size_t)a * lds_slots + s] = acc_identity(A.op[a]);
  }
  __syncthreads();
  const int rep = xcc_id() % g.nrep;

  const int64_t per_block = ceil_div_d(n, gridDim.x);
  const int64_t begin = (int64_t)blockIdx.x * per_block;
  const int64_t end = begin + per_block < n ? begin + per_block : n;
  bool out_of_range = false;
  for (int64_t r = begin + threadIdx.x; r < end; r += blockDim.x) {
    uint64_t k = group_key(keys, r, &out_of_range);
    uint64_t v[kMaxAggs];
    bool ok[kMaxAggs];
#pragma unroll
    for (int a = 0; a < (NACC > 0 ? NACC : kMaxAggs); ++a) {
      if (a < nacc) v[a] = acc_operand(vals, A.op[a], A.input[a], r, &ok[a]);
    }
    const int64_t grow = row_base + r;
    int lslot = -1;
    if (k != kEmptyKey) {
      uint32_t s = (uint32_t)fmix64(k) & lmask;
      for (int p = 0; p < kLdsProbes; ++p) {
        uint64_t cur = lkeys[s];
        if (cur == k) { lslot = (int)s; break; }
        if (cur == kEmptyKey) {
          unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long*>(&lkeys[s]), kEmptyKey, k);
          if (prev == kEmptyKey || prev == k) { lslot = (int)s; break; }
        }
        s = (s + 1) & lmask;
      }
    }
    if (lslot >= 0) {
#pragma unroll
      for (int a = 0; a < (NACC > 0 ? NACC : kMaxAggs); ++a) {
        if (a < nacc && ok[a]) acc_apply(&lacc[(size_t)a * lds_slots + lslot], A.op[a], v[a]);
      }
      if (grow < lfirst[lslot]) atomicMin(reinterpret_cast<long long*>(&lfirst[lslot]), (long long)grow);
    } else {
      int64_t gs = global_slot(g, rep, k);
      if (gs >= 0) {
#pragma unroll
        for (int a = 0; a < (NACC > 0 ? NACC : kMaxAggs); ++a) {
          if (a < nacc && ok[a]) acc_apply(gacc(g, rep, a, gs), A.op[a], v[a]);
        }
        atomicMin(reinterpret_cast<long long*>(gfirst(g, rep, gs)), (long long)grow);
      }
    }
  }
  if (out_of_range) *g.key_range = 1;
  __syncthreads();
  // merge the LDS table into the global table
  for (int s = threadIdx.x; s < lds_slots; s += blockDim.x) {
    uint64_t k = lkeys[s];
    if (k == kEmptyKey) continue;
    int64_t gs = global_slot(g, rep, k);
    if (gs < 0) continue;
    for (int a = 0; a < nacc; ++a) {
      uint64_t v = lacc[(size_t)a * lds_slots + s];
      if (v != acc_identity(A.op[a]) || A.op[a] == kAccSumF) acc_apply(gacc(g, rep, a, gs), A.op[a], v);
    }
    atomicMin(reinterpret_cast<long long*>(gfirst(g, rep, gs)), (long long)lfirst[s]);
  }
}

__global__ void groupby_init_kernel(GlobalTable g, AccSpec A) {
  int64_t total = g.stride * g.nrep;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    int rep = (int)(i / g.stride);
    int64_t s = i % g.stride;
    if (s < g.stride - 1) *gkey(g, rep, s) = kEmptyKey;
    *gfirst(g, rep, s) = LLONG_MAX;
    for (int a = 0; a < A.n; ++a) *gacc(g, rep, a, s) = acc_identity(A.op[a]);
  }
}

// Folds replicas 1.. into replica 0 and resets them to empty.
__global__ void groupby_merge_kernel(GlobalTable g, AccSpec A) {
  int64_t total = g.stride * (g.nrep - 1);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    int rep = 1 + (int)(i / g.stride);
    int64_t s = i % g.stride;
    int64_t f = *gfirst(g, rep, s);
    if (f == LLONG_MAX) continue;    // slot never used
    uint64_t k = (s == g.stride - 1) ? kEmptyKey : *gkey(g, rep, s);
    int64_t d = global_slot(g, 0, k);
    if (d >= 0) {
      for (int a = 0; a < A.n; ++a) {
        uint64_t v = *gacc(g, rep, a, s);
        if (v != acc_identity(A.op[a]) || A.op[a] == kAccSumF) acc_apply(gacc(g, 0, a, d), A.op[a], v);
        *gacc(g, rep, a, s) = acc_identity(A.op[a]);
      }
      atomicMin(reinterpret_cast<long long*>(gfirst(g, 0, d)), (long long)f);
    }
    if (s < g.stride - 1) *gkey(g, rep, s) = kEmptyKey;
    *gfirst(g, rep, s) = LLONG_MAX;
  }
}

// Marks the first row of every group in a bitmap over consumed rows.
__global__ void groupby_mark_kernel(GlobalTable g, unsigned long long* __restrict__ bits) {
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < g.stride;
       s += (int64_t)gridDim.x * blockDim.x) {
    int64_t f = g.first[s];
    if (f == LLONG_MAX) continue;
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

__global__ void groupby_emit_kernel(GlobalTable g, AccSpec A, const unsigned long long* __restrict__ bits,
                                    const uint32_t* __restrict__ word_prefix, GOutSpec o) {
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < g.stride;
       s += (int64_t)gridDim.x * blockDim.x) {
    int64_t f = g.first[s];
    if (f == LLONG_MAX) continue;
    uint64_t w = bits[f >> 6];
    uint64_t below = (f & 63) ? (w & (~0ULL >> (64 - (f & 63)))) : 0ULL;
    int64_t rank = (int64_t)word_prefix[f >> 6] + __popcll(below);
    uint64_t key = (s == g.stride - 1) ? kEmptyKey : g.keys[s];
    for (int j = 0; j < o.num_keys; ++j) {
      uint64_t v = ((key >> o.key_shift[j]) & key_field_mask(o.key_bits[j])) + (uint64_t)o.key_bias[j];
      store_w(o.key_values[j], o.key_w[j], rank + o.key_off[j], v);
    }
    for (int a = 0; a < o.num_aggs; ++a) {
      uint64_t v = g.acc[(size_t)a * g.stride + s];
      if (o.agg_is_f32[a]) {
        float fv = (float)__longlong_as_double((long long)v);
        v = (uint64_t)__float_as_uint(fv);
      }
      uint64_t cnt = o.agg_count_acc[a] >= 0 ? g.acc[(size_t)o.agg_count_acc[a] * g.stride + s] : 1;
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
// so one unsigned atomic min/max serves both.  out: [min, max, valid count].
__global__ void range_init_kernel(unsigned long long* out) {
  if (threadIdx.x == 0) {
    out[0] = ~0ULL;
    out[1] = 0ULL;
    out[2] = 0ULL;
  }
}

__global__ __launch_bounds__(256) void range_kernel(ColSet col, int64_t n, unsigned long long* out) {
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
  if ((threadIdx.x & 63) == 0 && cnt) {
    atomicMin(out + 0, lo);
    atomicMax(out + 1, hi);
    atomicAdd(out + 2, cnt);
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
  int64_t capacity = 0;
  uint64_t* keys = nullptr;
  uint64_t* accs = nullptr;
  int64_t* first = nullptr;
  int64_t rows = 0;
  unsigned long long* d_groups = nullptr;   // [0] groups, [1] overflow flag (int) + key range flag (int)
};

namespace fpdb {

static GlobalTable global_view(fpdb_group_table* t) {
  GlobalTable g;
  g.keys = t->keys;
  g.acc = t->accs;
  g.first = t->first;
  g.mask = (uint64_t)t->capacity - 1;
  g.stride = t->capacity + 1;
  g.nacc = t->acc.n;
  g.nrep = kReplicas;
  g.groups = t->d_groups;
  g.overflow = reinterpret_cast<int*>(t->d_groups + 1);
  g.key_range = reinterpret_cast<int*>(t->d_groups + 1) + 1;
  return g;
}

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
  FPDB_CHECK(max_groups >= 1, FPDB_ERR_INVALID, "group-by: max_groups must be positive");
  Context* c = as_ctx(ctx);
  auto* t = new fpdb_group_table();
  t->ctx = c;
  t->num_keys = num_keys;
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
  if (nacc > kMaxAggs) {
    delete t;
    return fail(FPDB_ERR_INVALID, "group-by: too many aggregates");
  }
  t->acc.n = nacc;
  int64_t cap = 64;
  while (cap < 2 * max_groups) cap <<= 1;
  t->capacity = cap;
  fpdb_status s = ctx_alloc(c, (uint64_t)cap * 8 * kReplicas, reinterpret_cast<void**>(&t->keys));
  if (s == FPDB_OK)
    s = ctx_alloc(c, (uint64_t)(cap + 1) * 8 * std::max(nacc, 1) * kReplicas, reinterpret_cast<void**>(&t->accs));
  if (s == FPDB_OK) s = ctx_alloc(c, (uint64_t)(cap + 1) * 8 * kReplicas, reinterpret_cast<void**>(&t->first));
  if (s == FPDB_OK) s = ctx_alloc(c, 16, reinterpret_cast<void**>(&t->d_groups));
  if (s == FPDB_OK) s = launch_fill(c, t->d_groups, 0, 16);
  if (s != FPDB_OK) {
    fpdb_groupby_destroy(t);
    return s;
  }
  GlobalTable g = global_view(t);
  hipLaunchKernelGGL(groupby_init_kernel, dim3(grid_for(c, (cap + 1) * kReplicas, 256)), dim3(256), 0, c->stream, g,
                     t->acc);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    fpdb_groupby_destroy(t);
    return fail(FPDB_ERR_HIP, hipGetErrorString(e));
  }
  *out = t;
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
  hipLaunchKernelGGL(range_init_kernel, dim3(1), dim3(64), 0, c->stream, d);
  if (num_rows > 0)
    hipLaunchKernelGGL(range_kernel, dim3(grid_for(c, num_rows, 256, 4)), dim3(256), 0, c->stream, cs, num_rows, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h, d, 24, hipMemcpyDeviceToHost, c->stream);
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
  if (t->keys) ctx_free(t->ctx, t->keys);
  if (t->accs) ctx_free(t->ctx, t->accs);
  if (t->first) ctx_free(t->ctx, t->first);
  if (t->d_groups) ctx_free(t->ctx, t->d_groups);
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
  int nacc = t->acc.n;
  // LDS table size: the largest power of two that fits the per-block budget
  size_t per_slot = 8 * (size_t)(2 + nacc);
  size_t budget = std::min<size_t>(c->max_lds_per_block, 128 * 1024);
  int slots = 256;
  while ((size_t)slots * 2 * per_slot <= budget) slots *= 2;
  size_t shmem = (size_t)slots * per_slot;
  int blocks = std::max(1, c->num_cus);
  int64_t max_blocks = ceil_div(n, 4096);
  blocks = (int)std::max<int64_t>(1, std::min<int64_t>(blocks, max_blocks));
  GlobalTable g = global_view(t);
  ProfScope ps(c, "groupby_consume");
  switch (nacc) {
#define FPDB_GB_CASE(N)                                                                                     \
  case N:                                                                                                   \
    hipLaunchKernelGGL(groupby_consume_kernel<N>, dim3(blocks), dim3(kGroupThreads), shmem, c->stream, k, vals, \
                       t->acc, n, t->rows, g, slots);                                                       \
    break;
    FPDB_GB_CASE(1)
    FPDB_GB_CASE(2)
    FPDB_GB_CASE(3)
    FPDB_GB_CASE(4)
    FPDB_GB_CASE(5)
#undef FPDB_GB_CASE
    default:
      hipLaunchKernelGGL(groupby_consume_kernel<0>, dim3(blocks), dim3(kGroupThreads), shmem, c->stream, k, vals,
                         t->acc, n, t->rows, g, slots);
      break;
  }
  FPDB_HIP(hipGetLastError());
  t->rows += n;
  return FPDB_OK;
}

fpdb_status fpdb_groupby_num_groups(fpdb_context* ctx, fpdb_group_table* t, int64_t* num_groups) {
  FPDB_CHECK(ctx && t && num_groups, FPDB_ERR_INVALID, "NULL argument");
  Context* c = as_ctx(ctx);
  if (t->rows > 0) {
    GlobalTable g = global_view(t);
    ProfScope ps(c, "groupby_merge");
    hipLaunchKernelGGL(groupby_merge_kernel, dim3(grid_for(c, g.stride * (kReplicas - 1), 256)), dim3(256), 0,
                       c->stream, g, t->acc);
    FPDB_HIP(hipGetLastError());
  }
  FPDB_HIP(hipMemcpyAsync(c->h_counter, t->d_groups, 16, hipMemcpyDeviceToHost, c->stream));
  int64_t sentinel_first = 0;
  FPDB_HIP(hipMemcpyAsync(&sentinel_first, t->first + t->capacity, 8, hipMemcpyDeviceToHost, c->stream));
  FPDB_HIP(hipStreamSynchronize(c->stream));
  FPDB_CHECK(reinterpret_cast<int*>(c->h_counter + 1)[1] == 0, FPDB_ERR_INVALID,
             "group-by: a key value lies outside the range the packed key was created for");
  FPDB_CHECK(reinterpret_cast<int*>(c->h_counter + 1)[0] == 0, FPDB_ERR_CAPACITY,
             "group-by: more groups than max_groups");
  *num_groups = c->h_counter[0] + (sentinel_first != LLONG_MAX ? 1 : 0);
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
    GlobalTable g = global_view(t);
    hipLaunchKernelGGL(groupby_mark_kernel, dim3(grid_for(c, g.stride, 256)), dim3(256), 0, c->stream, g, bits);
    hipLaunchKernelGGL(popcount_chunk_kernel, dim3((unsigned)chunks), dim3(1024), 0, c->stream, bits, words,
                       chunk_sums);
    hipLaunchKernelGGL(chunk_scan_kernel, dim3(1), dim3(1024), 0, c->stream, chunk_sums, chunks);
    hipLaunchKernelGGL(word_prefix_kernel, dim3((unsigned)chunks), dim3(1024), 0, c->stream, bits, words,
                       chunk_sums, word_prefix);
    hipLaunchKernelGGL(groupby_emit_kernel, dim3(grid_for(c, g.stride, 256)), dim3(256), 0, c->stream, g, t->acc,
                       bits, word_prefix, o);
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
