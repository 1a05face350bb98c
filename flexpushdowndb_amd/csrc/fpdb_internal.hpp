// Internal declarations shared by the HIP translation units of libfpdb_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/fpdb_amd.h"

namespace fpdb {

// ---- errors -------------------------------------------------------------

void set_error(const std::string& msg);
fpdb_status fail(fpdb_status code, const std::string& msg);

#define FPDB_HIP(call)                                                          \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess)                                                       \
      return ::fpdb::fail(FPDB_ERR_HIP, std::string(#call " failed: ") +        \
                                             hipGetErrorString(e_));            \
  } while (0)

#define FPDB_CHECK(cond, code, msg)                                             \
  do {                                                                          \
    if (!(cond)) return ::fpdb::fail((code), (msg));                            \
  } while (0)

#define FPDB_TRY(call)                                                          \
  do {                                                                          \
    fpdb_status s_ = (call);                                                    \
    if (s_ != FPDB_OK) return s_;                                               \
  } while (0)

// ---- context ---------------------------------------------------------------

struct ProfRecord {
  const char* name;
  hipEvent_t start, stop;
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  uint64_t budget = 0;              // 0 = unlimited
  uint64_t used = 0;                // bytes of live allocations (rounded sizes)
  std::unordered_map<void*, uint64_t> allocs;
  // freed blocks kept for reuse, by rounded size
  std::unordered_multimap<uint64_t, void*> cache;
  uint64_t cached = 0;
  std::mutex mu;
  // kernel timing: events recorded around named launches while enabled
  bool profiling = false;
  std::vector<ProfRecord> prof;
  std::vector<hipEvent_t> event_pool;
  int num_cus = 0;
  int num_xcds = 8;
  uint64_t hbm_bytes = 0;
  size_t max_lds_per_block = 64 * 1024;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  int64_t* d_counter = nullptr;     // small device scratch for counts (8 words)
  unsigned long long* d_partials = nullptr;   // per-workgroup partial counts (kMaxPartials)
  int64_t* h_counter = nullptr;     // pinned host mirror
  // read-back mailbox: pinned, coherent host memory the device writes
  // (word 0 a sequence number, words 1..7 values); see mail_read
  uint64_t* h_mail = nullptr;
  uint64_t* d_mail = nullptr;       // the same memory, device view
  uint64_t mail_seq = 0;
  // Device address of this device's copy of the library's static no-allocation
  // scratch array (kNoAllocScratchBytes), looked up once per device when the
  // first context on it is created.  Shared by every context on the device.
  uint8_t* noalloc_scratch = nullptr;
};

// The static scratch array of the no-allocation entry points: tile counts and
// tile bases of a compaction of at most kNoAllocMaxRows rows.
constexpr int64_t kNoAllocMaxRows = int64_t(1) << 28;
constexpr int64_t kNoAllocMaxTiles = kNoAllocMaxRows / 4096;
constexpr uint64_t kNoAllocScratchBytes = (uint64_t)kNoAllocMaxTiles * 12;

fpdb_status ctx_alloc(Context* c, uint64_t bytes, void** out);
fpdb_status ctx_free(Context* c, void* p);

// Brackets one kernel launch with events on the context stream when the
// context's profiling mode is on.
hipEvent_t prof_event(Context* c);
struct ProfScope {
  Context* c;
  const char* name;
  hipEvent_t start = nullptr;
  ProfScope(Context* ctx, const char* n) : c(ctx), name(n) {
    if (c->profiling) {
      start = prof_event(c);
      if (start) (void)hipEventRecord(start, c->stream);
    }
  }
  ~ProfScope() {
    if (c->profiling && start) {
      hipEvent_t stop = prof_event(c);
      if (stop) {
        (void)hipEventRecord(stop, c->stream);
        c->prof.push_back(ProfRecord{name, start, stop});
      }
    }
  }
};
// Per-workgroup partial counts.  A same-address device atomic serializes at
// about 12 ns each on gfx950, so kernels that count (selected rows, inserted
// keys) store one partial per workgroup and a one-workgroup kernel adds the
// partials into the destination (stored, not added).
constexpr int kMaxPartials = 1 << 14;
fpdb_status launch_sum_partials(Context* c, int blocks, int64_t* dst);
fpdb_status launch_fill(Context* c, void* d, uint8_t v, uint64_t bytes);
// Kernel-side failure flag (a bounded wait gave up, a table overflowed); zero
// unless a kernel reported a failure.
inline int* ctx_err_flag(Context* c) { return reinterpret_cast<int*>(c->d_counter + 1); }
// Synchronizes, reads the count a kernel left in d_counter[0] and checks (and
// clears) the failure flag.
fpdb_status read_count(Context* c, int64_t* count);
// Reads nwords (<= 7) 8-byte words from device memory once the work queued on
// the context stream before it has run: a one-wave kernel stores them and then
// a sequence number into pinned host memory, and the host spins on the
// sequence number -- about the latency of the kernel itself, where a copy and
// a stream synchronize wait for the host thread to be woken.  After 50 ms of
// spinning it falls back to synchronizing the stream (which reports a failed
// kernel).
fpdb_status mail_read(Context* c, const void* d_src, int nwords, uint64_t* out);

inline Context* as_ctx(fpdb_context* c) { return reinterpret_cast<Context*>(c); }

// ---- column helpers ----------------------------------------------------------

int type_width(int32_t type);      // bytes, 0 if unknown
bool type_is_float(int32_t type);
bool type_is_signed_int(int32_t type);

constexpr int kMaxCols = 24;

// Kernel-argument view of a set of columns.
struct ColSet {
  const uint8_t* values[kMaxCols];
  const uint8_t* validity[kMaxCols];
  int64_t offset[kMaxCols];
  int32_t width[kMaxCols];
  int32_t type[kMaxCols];
  int32_t n;
};

struct OutColSet {
  uint8_t* values[kMaxCols];
  uint8_t* validity[kMaxCols];
  int64_t offset[kMaxCols];
  int32_t width[kMaxCols];
  int32_t n;
  int64_t capacity;   // rows every output column has room for (INT64_MAX with no columns)
};

fpdb_status make_colset(const fpdb_column* cols, int n, int64_t num_rows, ColSet* out,
                        const char* what);

// utf8 columns (strings.hip).  utf8_gather: rows d_rows[0, n) of `in` (-1 =
// null) into `out` from row 0; FPDB_ERR_CAPACITY if the bytes pass
// out.data_bytes (synchronizes).  utf8_copy_rows: rows [0, n) of `in` after
// output row out_row, whose offsets up to out_row are written (synchronizes).
fpdb_status utf8_gather(Context* c, const fpdb_column& in, const int64_t* d_rows, int64_t n, const fpdb_column& out);
fpdb_status utf8_copy_rows(Context* c, const fpdb_column& in, int64_t n, const fpdb_column& out, int64_t out_row);

// Splits an input/output column list into its fixed-width columns (in order)
// and the positions of its utf8 columns.
inline fpdb_status split_utf8(const fpdb_column* in, const fpdb_column* out, int n, std::vector<fpdb_column>* fin,
                              std::vector<fpdb_column>* fout, std::vector<int>* utf8, const char* what) {
  FPDB_CHECK(n >= 0 && (n == 0 || (in && out)), FPDB_ERR_INVALID, std::string(what) + ": NULL columns");
  for (int i = 0; i < n; ++i) {
    if (in[i].type == FPDB_TYPE_UTF8 || out[i].type == FPDB_TYPE_UTF8) {
      FPDB_CHECK(in[i].type == out[i].type, FPDB_ERR_INVALID, std::string(what) + ": utf8 column into another type");
      utf8->push_back(i);
    } else {
      fin->push_back(in[i]);
      fout->push_back(out[i]);
    }
  }
  return FPDB_OK;
}
fpdb_status make_outcolset(const fpdb_column* cols, int n, OutColSet* out, const char* what);

// ---- device helpers ------------------------------------------------------------

__device__ __forceinline__ uint64_t load_zext(const uint8_t* base, int width, int64_t i) {
  switch (width) {
    case 8: return reinterpret_cast<const uint64_t*>(base)[i];
    case 4: return reinterpret_cast<const uint32_t*>(base)[i];
    case 2: return reinterpret_cast<const uint16_t*>(base)[i];
    default: return base[i];
  }
}

// Sign-extending load used for join/group keys of signed integer types.
__device__ __forceinline__ int64_t load_sext(const uint8_t* base, int width, int64_t i) {
  switch (width) {
    case 8: return reinterpret_cast<const int64_t*>(base)[i];
    case 4: return reinterpret_cast<const int32_t*>(base)[i];
    case 2: return reinterpret_cast<const int16_t*>(base)[i];
    default: return reinterpret_cast<const int8_t*>(base)[i];
  }
}

__device__ __forceinline__ void store_w(uint8_t* base, int width, int64_t i, uint64_t v) {
  switch (width) {
    case 8: reinterpret_cast<uint64_t*>(base)[i] = v; break;
    case 4: reinterpret_cast<uint32_t*>(base)[i] = (uint32_t)v; break;
    case 2: reinterpret_cast<uint16_t*>(base)[i] = (uint16_t)v; break;
    default: base[i] = (uint8_t)v; break;
  }
}

// store_w with non-temporal 8- and 4-byte stores: for output columns written
// once, which then do not push gathered or reused lines out of the L2.
__device__ __forceinline__ void store_w_nt(uint8_t* base, int width, int64_t i, uint64_t v) {
  switch (width) {
    case 8: __builtin_nontemporal_store(v, reinterpret_cast<uint64_t*>(base) + i); break;
    case 4: __builtin_nontemporal_store((uint32_t)v, reinterpret_cast<uint32_t*>(base) + i); break;
    default: store_w(base, width, i, v); break;
  }
}

// Copies count 8-byte words src[i * src_stride] into dst[i] (LDS) with the
// whole workgroup.  Each thread issues kStageWords loads before its first LDS
// write: a loop of one load and one write per trip would wait for memory once
// per trip.  The caller synchronises the workgroup before reading dst.
constexpr int kStageWords = 8;
__device__ __forceinline__ void stage_words(uint64_t* dst, const uint64_t* __restrict__ src, uint64_t count,
                                            uint64_t src_stride = 1) {
  const uint64_t nt = blockDim.x;
  for (uint64_t i0 = threadIdx.x; i0 < count; i0 += nt * kStageWords) {
    uint64_t v[kStageWords];
#pragma unroll
    for (int j = 0; j < kStageWords; ++j) {
      const uint64_t i = i0 + (uint64_t)j * nt;
      v[j] = src[(i < count ? i : i0) * src_stride];
    }
#pragma unroll
    for (int j = 0; j < kStageWords; ++j) {
      const uint64_t i = i0 + (uint64_t)j * nt;
      if (i < count) dst[i] = v[j];
    }
  }
}

// Loads N values of one column at once (indices already clamped to valid
// rows): the width switch sits outside the per-value loop, so the N loads are
// issued back to back instead of each in a branch that waits for it.
template <int N>
__device__ __forceinline__ void load_zext_n(const uint8_t* base, int width, const int64_t (&idx)[N],
                                            uint64_t (&out)[N]) {
  switch (width) {
    case 8: {
      const uint64_t* p = reinterpret_cast<const uint64_t*>(base);
#pragma unroll
      for (int i = 0; i < N; ++i) out[i] = p[idx[i]];
      break;
    }
    case 4: {
      const uint32_t* p = reinterpret_cast<const uint32_t*>(base);
#pragma unroll
      for (int i = 0; i < N; ++i) out[i] = p[idx[i]];
      break;
    }
    case 2: {
      const uint16_t* p = reinterpret_cast<const uint16_t*>(base);
#pragma unroll
      for (int i = 0; i < N; ++i) out[i] = p[idx[i]];
      break;
    }
    default: {
#pragma unroll
      for (int i = 0; i < N; ++i) out[i] = base[idx[i]];
      break;
    }
  }
}

// load_zext_n with non-temporal loads for 4- and 8-byte columns: for columns
// that stream through a kernel once, so they do not push out of the L2 what
// the kernel gathers at random.
template <int N>
__device__ __forceinline__ void load_zext_n_nt(const uint8_t* base, int width, const int64_t (&idx)[N],
                                               uint64_t (&out)[N]) {
  switch (width) {
    case 8: {
      const uint64_t* p = reinterpret_cast<const uint64_t*>(base);
#pragma unroll
      for (int i = 0; i < N; ++i) out[i] = __builtin_nontemporal_load(p + idx[i]);
      break;
    }
    case 4: {
      const uint32_t* p = reinterpret_cast<const uint32_t*>(base);
#pragma unroll
      for (int i = 0; i < N; ++i) out[i] = __builtin_nontemporal_load(p + idx[i]);
      break;
    }
    default: load_zext_n<N>(base, width, idx, out); break;
  }
}

__device__ __forceinline__ bool valid_bit(const uint8_t* validity, int64_t i) {
  return validity == nullptr || ((validity[i >> 3] >> (i & 7)) & 1);
}

// Arrow 6 Hashing::helper_8B for a key of <= 8 bytes, zero-extended:
// (uint32) bswap64(x * 0xC2B2AE3D27D4EB4F) == bswap32(high 32 bits).
__host__ __device__ __forceinline__ uint32_t arrow6_hash_word(uint64_t x) {
  uint64_t p = x * 0xC2B2AE3D27D4EB4FULL;
  uint32_t hi = (uint32_t)(p >> 32);
  return (hi >> 24) | ((hi >> 8) & 0xff00u) | ((hi << 8) & 0xff0000u) | (hi << 24);
}

// Arrow 6 CombineHashesImp.
__host__ __device__ __forceinline__ uint32_t arrow6_combine(uint32_t prev, uint32_t h) {
  return prev ^ (h + 0x9e3779b9u + (prev << 6) + (prev >> 2));
}

__device__ __forceinline__ uint32_t hash_row(const ColSet& k, int64_t row) {
  uint32_t h = 0;
  for (int c = 0; c < k.n; ++c) {
    int64_t i = row + k.offset[c];
    uint32_t hc = arrow6_hash_word(load_zext(k.values[c], k.width[c], i));
    if (!valid_bit(k.validity[c], i)) hc = 0;
    h = (c == 0) ? hc : arrow6_combine(h, hc);
  }
  return h;
}

// ---------------------------------------------------------------------------
// key access modes: 0 = generic column set, 1 = one 8-byte column without
// nulls, 2 = one 4-byte column without nulls.
// ---------------------------------------------------------------------------
template <int MODE>
__device__ __forceinline__ uint32_t key_hash(const ColSet& k, int64_t row) {
  if constexpr (MODE == 1) {
    return arrow6_hash_word(reinterpret_cast<const uint64_t*>(k.values[0])[row + k.offset[0]]);
  } else if constexpr (MODE == 2) {
    return arrow6_hash_word(reinterpret_cast<const uint32_t*>(k.values[0])[row + k.offset[0]]);
  } else {
    return hash_row(k, row);
  }
}

inline int key_mode(const ColSet& k) {
  if (k.n == 1 && k.validity[0] == nullptr) {
    if (k.width[0] == 8) return 1;
    if (k.width[0] == 4) return 2;
  }
  return 0;
}

// 64-bit finaliser (Murmur3 fmix64) used for the join and group tables.
__host__ __device__ __forceinline__ uint64_t fmix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdULL;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ULL;
  k ^= k >> 33;
  return k;
}

// Stores this workgroup's count (the sum of every thread's `mine`) into
// partials[blockIdx.x]; every thread of the workgroup must call it.
__device__ __forceinline__ void block_count_store(unsigned long long* partials, unsigned long long mine) {
  __shared__ unsigned long long s_block_count;
  if (threadIdx.x == 0) s_block_count = 0;
  __syncthreads();
  for (int off = 32; off > 0; off >>= 1) mine += __shfl_xor(mine, off);
  if ((threadIdx.x & 63) == 0 && mine) atomicAdd(&s_block_count, mine);
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = s_block_count;
}

// Wave-wide inclusive prefix sum of a 32-bit value.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t t = __shfl_up(v, off);
    if (lane >= off) v += t;
  }
  return v;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace fpdb
