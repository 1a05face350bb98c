// Star join with full predicate transfer over a partitioned fact table
// (SmallToLargePredTransOrder, fpdb-executor/src/physical/transform/pred-trans/
// SmallToLargePredTransOrder.cpp:38-214, with the partitioned semantics the
// reference lacks):
//   1. each dimension is filtered (attr <op> constant) and a forward blocked
//      bloom filter is built over the kept keys;
//   2. every fact partition is filtered by all forward filters;
//   3. backward filters are built over the union of the surviving fact rows'
//      foreign keys (the OR of per-partition filters) and applied to the kept
//      dimension rows;
//   4. the survivors are joined with the reduced dimensions (hash tables with
//      the group attribute inline) and grouped on two dimension attributes
//      (COUNT(*), SUM(measure)); groups are numbered by first appearance in
//      fact order, so per-partition partials merge in partition order.
//
// Two executions of the same plan:
//   plain    per-partition launches of the operator entry points (bloom filter
//            use, hash-join probe, group-by), as the drop-in operators run it;
//   batched  one persistent launch: workgroups stay resident and pass grid
//            barriers between the phases, so the whole predicate-transfer
//            chain, the joins and the aggregation are a single kernel; it can
//            be captured in a hipGraph and replayed.
#include "fpdb_internal.hpp"
#include "compact.hpp"

#include <algorithm>
#include <climits>
#include <vector>

namespace fpdb {

constexpr int kStarMaxDims = 8;
constexpr int kStarStateWords = 64 + kStarMaxDims;   // the state words of a star join (see StarDev::state)
constexpr int kStarInitThreads = 256;
constexpr int kStarInitRows = 8;          // dimension rows a thread per trip in the init kernel
constexpr int kStarInitMaxBlocks = 2048;  // init workgroups at most (its partial counts: 2 a persistent thread)
static_assert(kStarInitMaxBlocks % 1024 == 0, "whole loads of partial counts per persistent thread");
constexpr uint32_t kStarEmptyRow = 0xFFFFFFFFu;
constexpr uint64_t kStarEmptyKey = ~0ULL;

struct alignas(16) StarSlot {
  int64_t key;
  uint32_t row;
  uint32_t payload;
};

// Device view of one dimension.
struct StarDimDev {
  const int64_t* key;
  const int32_t* attr;
  const int32_t* gattr;      // group attribute, or nullptr
  int64_t rows;
  int64_t cst;
  int32_t op;
  int32_t ht_shift;          // hash-table capacity >= kept rows << ht_shift
  uint8_t* kept;             // rows bytes
  uint64_t* fbf;             // forward filter blocks (capacity for `rows` keys)
  int32_t fbf_lnb;           // log2 of that capacity in blocks (the init kernel zeroes all of it)
  uint64_t* bbf;             // backward filter blocks (capacity for all fact rows)
  StarSlot* ht;              // hash table storage (ht_mask + 1 slots); a run uses the
  uint64_t ht_mask;          // first star_ht_mask(...) + 1 of them
};

// A run of fact rows in one set of buffers: the whole fact table, or one
// partition held in buffers of its own.
struct StarSeg {
  int64_t base;              // global number of its first row
  int64_t rows;
  const int64_t* fk[kStarMaxDims];
  const int64_t* meas;
};

struct StarDev {
  StarDimDev dim[kStarMaxDims];
  int32_t ndims;
  int32_t order[kStarMaxDims];     // forward filter order over the fact columns
  int32_t gdim[2];                 // dimensions whose group attribute forms the key
  const StarSeg* segs;             // device: the fact rows' segments, in row order
  int32_t nsegs;
  int64_t fact_rows;
  // survivors of the forward pass
  int64_t* s_idx;                  // original fact row
  int64_t* s_fk;                   // ndims x fact_rows
  int64_t* s_meas;
  // state words: [0] survivor count, [1..ndims] kept count per dim,
  // [16..31] barrier, [32] error flag (all reset by the init kernel);
  // [64 + d] log2 of the backward filter blocks of dim d that may be set,
  // kept across runs so that the next init kernel clears them
  unsigned long long* state;
  // group table
  uint64_t* g_keys;
  unsigned long long* g_cnt;
  unsigned long long* g_sum;
  long long* g_first;
  uint64_t g_mask;
  int32_t tile_rows;
  uint64_t barrier_ticks;          // grid barrier wait bound, device wall-clock ticks
  int32_t bbf_cap_lnb;             // backward filter capacity, log2 blocks
  // kept rows per dimension per init workgroup ([kStarMaxDims][init_blocks]),
  // written by the init kernel and summed by every persistent workgroup
  unsigned long long* kept_part;
  int32_t init_blocks;
};

__device__ __forceinline__ int log_blocks_for(int64_t n) {
  int64_t desired = n * 8 < 512 ? 512 : n * 8;
  int lb = 0;
  while ((int64_t(1) << lb) < desired) ++lb;
  return lb - 6;
}

__device__ __forceinline__ bool star_cmp(int32_t v, int32_t op, int64_t c) {
  switch (op) {
    case FPDB_CMP_LT: return v < c;
    case FPDB_CMP_LE: return v <= c;
    case FPDB_CMP_GT: return v > c;
    case FPDB_CMP_GE: return v >= c;
    case FPDB_CMP_EQ: return v == c;
    default: return v != c;
  }
}

__device__ __forceinline__ uint64_t rotl64s(uint64_t x, int n) { return (x << n) | (x >> ((-n) & 63)); }

__device__ __forceinline__ void bf_insert(uint64_t* blocks, int lnb, const uint64_t* s_masks, int64_t key) {
  uint32_t h = arrow6_hash_word((uint64_t)key);
  uint64_t m = rotl64s(s_masks[h & (kMaskNum - 1)], (h >> kMaskLogNum) & 63);
  uint64_t b = ((uint64_t)h >> (kMaskLogNum + 6)) & ((1ULL << lnb) - 1);
  __hip_atomic_fetch_or(&blocks[b], m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// LDS for the forward filters of the fact pass (filters that do not fit are
// read from HBM / L2), and fact rows per thread per step.
constexpr int kStarFilterLdsWords = 8192;   // 64 KiB
constexpr size_t kStarDynamicLds = kStarFilterLdsWords * 8;
constexpr int kStarItems = 8;      // default fact rows per thread per step (tuning: 4, 8 or 16)
constexpr int kStarThreads = 1024;   // threads per workgroup of the persistent kernel
constexpr uint64_t kStarBarrierTimeoutUs = 200000;   // grid barrier wait bound (a normal wait is < 1 ms)

__device__ __forceinline__ bool bf_find(const uint64_t* blocks, int lnb, const uint64_t* s_masks, int64_t key) {
  uint32_t h = arrow6_hash_word((uint64_t)key);
  uint64_t m = rotl64s(s_masks[h & (kMaskNum - 1)], (h >> kMaskLogNum) & 63);
  uint64_t b = ((uint64_t)h >> (kMaskLogNum + 6)) & ((1ULL << lnb) - 1);
  return (blocks[b] & m) == m;
}

// Grid barrier for a launch whose workgroups are all resident (the grid is
// sized from the occupancy calculator and checked against it at every
// launch).  Sense-free: a generation counter advances when the last workgroup
// arrives.  The wait is bounded by wall-clock time (P.barrier_ticks of the
// device's constant-rate clock): a workgroup that waits longer sets the error
// word.  Once the error word is set every workgroup stops at its next barrier
// (returns false, and the kernel returns), so a missed barrier ends the launch
// within about one bound, with FPDB_ERR_INTERNAL reported by the result call.
// Only thread 0 fences, once each way: __syncthreads orders the workgroup's
// stores before its agent-scope release fence (an L2 write-back on this part)
// and its acquire fence (an L2 invalidate) before the workgroup's later loads.
// The barrier words themselves are relaxed atomics (they go past the L2s, and
// a fence per access would write back or invalidate the L2 each time: with
// acquire/release on every access a barrier of 512 workgroups took ~39 us,
// with one fence each way ~13 us).  Each arrival is a relaxed increment after
// the release fence; the last arriver acquires (so every workgroup's stores
// happen before what follows), resets the count and bumps the generation with
// a release; the others poll the generation and then acquire.  The count and
// the generation sit in different 128-byte lines.
__device__ __forceinline__ bool star_barrier(const StarDev& P) {
  unsigned long long* count = P.state + 16;
  unsigned long long* gen = P.state + 48;
  unsigned long long* err = P.state + 32;
  __shared__ int s_ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    int ok = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    if (ok) {
      const unsigned long long g = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned long long arrived =
          __hip_atomic_fetch_add(count, 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
      if (arrived == gridDim.x) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __hip_atomic_store(count, 0ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(gen, 1ULL, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        const uint64_t t0 = wall_clock64();
        while (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
          __builtin_amdgcn_s_sleep(2);
          if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
            ok = 0;   // another workgroup gave up
            break;
          }
          if (wall_clock64() - t0 > P.barrier_ticks) {
            __hip_atomic_store(err, 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ok = 0;
            break;
          }
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // the workgroup observes the other workgroups' writes
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

// Hash-table mask of a dimension for this run: capacity the smallest power of
// two >= its kept rows << ht_shift (at least 64), within the storage.
// Slots of a dimension's hash table in use for `kept` filtered rows.
__device__ __forceinline__ uint64_t star_ht_mask(const StarDimDev& D, uint64_t kept) {
  uint64_t cap = 64;
  while (cap < (kept << D.ht_shift)) cap <<= 1;
  return cap - 1 < D.ht_mask ? cap - 1 : D.ht_mask;
}

__device__ __forceinline__ void ht_insert(StarSlot* ht, uint64_t mask, int64_t key, uint32_t row, uint32_t payload) {
  uint64_t s = fmix64((uint64_t)key) & mask;
  for (uint64_t p = 0; p <= mask; ++p) {
    uint32_t old = atomicCAS(&ht[s].row, kStarEmptyRow, row);
    if (old == kStarEmptyRow) {
      ht[s].key = key;
      ht[s].payload = payload;
      return;
    }
    s = (s + 1) & mask;
  }
}

// Returns true and the payload if key is present (keys are unique per dimension).
__device__ __forceinline__ bool ht_find(const StarSlot* ht, uint64_t mask, int64_t key, uint32_t* payload) {
  uint64_t s = fmix64((uint64_t)key) & mask;
  for (uint64_t p = 0; p <= mask; ++p) {   // at most every slot once, even in a table left full
    uint4 v = reinterpret_cast<const uint4*>(ht)[s];
    if (v.z == kStarEmptyRow) return false;
    int64_t k = (int64_t)((uint64_t)v.x | ((uint64_t)v.y << 32));
    if (k == key) {
      *payload = v.w;
      return true;
    }
    s = (s + 1) & mask;
  }
  return false;
}

// Group slots 0..g_mask hold keys; slot g_mask + 1 is reserved for the one
// key equal to the empty marker.
__device__ __forceinline__ void group_add(const StarDev& P, uint64_t gkey, int64_t meas, int64_t first) {
  if (gkey == kStarEmptyKey) {
    uint64_t s = P.g_mask + 1;
    atomicAdd(&P.g_cnt[s], 1ULL);
    atomicAdd(&P.g_sum[s], (unsigned long long)meas);
    atomicMin(&P.g_first[s], (long long)first);
    return;
  }
  uint64_t s = fmix64(gkey) & P.g_mask;
  for (uint64_t p = 0; p <= P.g_mask; ++p) {
    uint64_t cur = __hip_atomic_load(&P.g_keys[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == kStarEmptyKey) {
      unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long*>(&P.g_keys[s]), kStarEmptyKey, gkey);
      if (prev == kStarEmptyKey || prev == gkey) cur = gkey;
      else cur = prev;
    }
    if (cur == gkey) {
      atomicAdd(&P.g_cnt[s], 1ULL);
      atomicAdd(&P.g_sum[s], (unsigned long long)meas);
      atomicMin(&P.g_first[s], (long long)first);
      return;
    }
    s = (s + 1) & P.g_mask;
  }
  __hip_atomic_store(P.state + 32, 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- phases shared by the persistent kernel and the plain path ---------------

// The fact segments' column pointers are read from device memory (the
// segment table), so the compiler cannot tell that they point to global
// memory and would load through them with flat instructions.  Flat loads
// also count against the LDS wait counter: each wait for a filter word or mask
// in LDS would then wait for every key loaded ahead as well.  Loads through
// this view are global loads.
typedef const int64_t __attribute__((address_space(1))) GlobalI64;
__device__ __forceinline__ GlobalI64* global_view(const int64_t* p) { return (GlobalI64*)p; }

// Orders one wave's LDS writes before its other lanes' reads (and reads
// before later writes) of the same words.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Forward-pass survivors are appended with one atomic per wavefront.
__device__ __forceinline__ void append_survivor(const StarDev& P, const StarSeg& S, bool pass, int64_t lr) {
  uint64_t mask = __ballot(pass);
  if (!mask) return;
  const int lane = threadIdx.x & 63;
  int leader = __builtin_ctzll(mask);
  unsigned long long base = 0;
  if (lane == leader)
    base = __hip_atomic_fetch_add(P.state + 0, (unsigned long long)__popcll(mask), __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
  base = __shfl(base, leader);
  if (pass) {
    uint64_t below = lane ? (mask & (~0ULL >> (64 - lane))) : 0ULL;
    int64_t pos = (int64_t)(base + __popcll(below));
    P.s_idx[pos] = S.base + lr;
    for (int d = 0; d < P.ndims; ++d) P.s_fk[(int64_t)d * P.fact_rows + pos] = global_view(S.fk[d])[lr];
    P.s_meas[pos] = global_view(S.meas)[lr];
  }
}

template <int ITEMS>
__global__ __launch_bounds__(kStarThreads) void star_persistent_kernel(StarDev P) {
  __shared__ uint64_t s_masks[kMaskNum];
  // per wave: the rows of one step that pass the first filter, on top of
  // fewer than 64 left from the steps before
  constexpr int kStackCap = 64 * (ITEMS + 1);
  __shared__ uint32_t s_stack[(kStarThreads / 64) * kStackCap];
  for (int i = threadIdx.x; i < kMaskNum; i += blockDim.x) s_masks[i] = c_masks.m[i];
  __syncthreads();
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gthreads = (int64_t)gridDim.x * blockDim.x;

  // A1 ran in the init kernel (kept flags and per-workgroup counts: a launch
  // boundary instead of a grid barrier).  Every workgroup sums the counts
  // into LDS once: the later phases size the filters and hash tables from
  // them (device-scope loads go to memory, and one per dimension per phase
  // in sequence added up to microseconds).
  __shared__ unsigned long long s_kept_rows[kStarMaxDims];
  if (threadIdx.x < kStarMaxDims) s_kept_rows[threadIdx.x] = 0;
  __syncthreads();
  {
    const int G = P.init_blocks;
    unsigned long long v[kStarMaxDims];
#pragma unroll
    for (int d = 0; d < kStarMaxDims; ++d) {
      v[d] = 0;
#pragma unroll
      for (int q = 0; q < kStarInitMaxBlocks / kStarThreads; ++q) {
        const int b = threadIdx.x + q * kStarThreads;
        if (d < P.ndims && b < G) v[d] += P.kept_part[(int64_t)d * G + b];
      }
    }
#pragma unroll
    for (int d = 0; d < kStarMaxDims; ++d) {
      if (d >= P.ndims) break;
      unsigned long long x = v[d];
      for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
      if ((threadIdx.x & 63) == 0 && x) atomicAdd(&s_kept_rows[d], x);
    }
  }
  __syncthreads();
  // the counts also go where the filter inspection reads them
  if (blockIdx.x == 0 && threadIdx.x < P.ndims) P.state[1 + threadIdx.x] = s_kept_rows[threadIdx.x];
  // A2: insert the kept keys into the forward filters (the init kernel zeroed
  // them at capacity, so no barrier separates a clear from the inserts), and
  // clear the used part of the hash tables (first used in phase C)
  for (int d = 0; d < P.ndims; ++d) {
    const StarDimDev& D = P.dim[d];
    int lnb = log_blocks_for((int64_t)s_kept_rows[d]);
    for (int64_t r = gtid; r < D.rows; r += gthreads)
      if (D.kept[r]) bf_insert(D.fbf, lnb, s_masks, D.key[r]);
    const uint4 empty = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
    const int64_t hm = (int64_t)star_ht_mask(D, s_kept_rows[d]);
    for (int64_t s = gtid; s <= hm; s += gthreads) reinterpret_cast<uint4*>(D.ht)[s] = empty;
  }
  if (!star_barrier(P)) return;
  // B: forward pass over every fact row (all partitions), filters in `order`.
  // The filters that fit (in that order) are copied into LDS; each thread
  // takes ITEMS rows per step with the first filter's keys loaded
  // together, and only rows passing the first filter read the next filters'
  // keys, a wave of such rows at a time.
  {
    // per-dimension filter size and LDS offset (-1: read from global memory),
    // in LDS: they are indexed by runtime dimension numbers
    extern __shared__ uint64_t s_filt[];
    __shared__ int lnb[kStarMaxDims];
    __shared__ int64_t loff[kStarMaxDims];
    int64_t used = 0;
    for (int i = 0; i < P.ndims; ++i) {
      int d = P.order[i];
      const int lb = log_blocks_for((int64_t)s_kept_rows[d]);
      int64_t words = int64_t(1) << lb;
      const bool fits = used + words <= kStarFilterLdsWords;
      if (threadIdx.x == 0) {
        lnb[d] = lb;
        loff[d] = fits ? used : -1;
      }
      if (fits) {
        stage_words(s_filt + used, P.dim[d].fbf, (uint64_t)words);
        used += words;
      }
    }
    __syncthreads();
    const int d0 = P.order[0];
    const int d1 = P.ndims > 1 ? P.order[1] : d0;
    // the first two filters' geometry, wave-uniform, in registers
    const int lnb0 = __builtin_amdgcn_readfirstlane(lnb[d0]);
    const int lnb1 = __builtin_amdgcn_readfirstlane(lnb[d1]);
    const bool lds0 = loff[d0] >= 0, lds1 = loff[d1] >= 0;
    const uint64_t* lf0 = s_filt + (lds0 ? loff[d0] : 0);
    const uint64_t* lf1 = s_filt + (lds1 ? loff[d1] : 0);
    const uint64_t* gf0 = P.dim[d0].fbf;
    const uint64_t* gf1 = P.dim[d1].fbf;
    const int64_t step = gthreads * ITEMS;
    // Rows that pass the first filter are pushed onto a per-wave stack in
    // LDS; whole waves of them (64 rows) are popped and run through the other
    // filters, so those filters cost one test per passing row instead of one
    // per row.  Survivor order does not matter: a group's first row is the
    // minimum of its rows' global numbers.
    const int lane = threadIdx.x & 63;
    uint32_t* stack = s_stack + (threadIdx.x >> 6) * kStackCap;
    // the filters after the first, for one row a lane (a wave at a time)
    auto rest = [&](const StarSeg& S, bool live, uint32_t lr) {
      bool pass = live;
      if (P.ndims > 1) {
        const int64_t k = global_view(S.fk[d1])[live ? lr : 0u];
        const bool hit = lds1 ? bf_find(lf1, lnb1, s_masks, k) : bf_find(gf1, lnb1, s_masks, k);
        pass = pass && hit;
      }
      for (int i = 2; i < P.ndims; ++i) {
        if (!__ballot(pass)) break;   // no row left in the wave
        const int d = P.order[i];
        const int64_t k = global_view(S.fk[d])[pass ? lr : 0u];
        const bool hit = loff[d] >= 0 ? bf_find(s_filt + loff[d], lnb[d], s_masks, k)
                                      : bf_find(P.dim[d].fbf, lnb[d], s_masks, k);
        pass = pass && hit;
      }
      append_survivor(P, S, pass, (int64_t)lr);
    };
    // segment by segment (one for a contiguous fact table, else one per
    // partition), rows numbered within the segment
    for (int sg = 0; sg < P.nsegs; ++sg) {
    const StarSeg& S = P.segs[sg];   // read in place: a copy indexed by runtime dimension would live in scratch
    if (S.rows == 0) continue;
    const int64_t rounded = (S.rows + step - 1) / step * step;   // whole waves iterate together
    const int64_t last = S.rows - 1;
    GlobalI64* col0 = global_view(S.fk[d0]);
    // each step's first-filter keys are loaded during the step before
    int64_t nk[ITEMS];
    {
      const int64_t q0 = (int64_t)blockIdx.x * blockDim.x * ITEMS + threadIdx.x;
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) {
        const int64_t r = q0 + (int64_t)j * blockDim.x;
        nk[j] = __builtin_nontemporal_load(col0 + (r <= last ? r : last));   // streamed once
      }
    }
    int qn = 0;   // rows on this wave's stack (wave-uniform, < 64 between steps)
    for (int64_t r0 = (int64_t)blockIdx.x * blockDim.x * ITEMS + threadIdx.x; r0 < rounded; r0 += step) {
      int64_t fk[ITEMS];
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) fk[j] = nk[j];
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) {
        const int64_t r = r0 + step + (int64_t)j * blockDim.x;
        nk[j] = __builtin_nontemporal_load(col0 + (r <= last ? r : last));
      }
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) {
        const int64_t r = r0 + (int64_t)j * blockDim.x;
        const bool hit = lds0 ? bf_find(lf0, lnb0, s_masks, fk[j]) : bf_find(gf0, lnb0, s_masks, fk[j]);
        const bool pass = r <= last && hit;
        const uint64_t m = __ballot(pass);
        const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (pass) stack[qn + (int)below] = (uint32_t)r;
        qn += __popcll(m);
      }
      while (qn >= 64) {
        wave_lds_sync();
        const uint32_t lr = stack[qn - 64 + lane];
        qn -= 64;
        rest(S, true, lr);
        wave_lds_sync();   // the next pushes overwrite the popped entries
      }
    }
    if (qn > 0) {   // the segment's last passing rows
      wave_lds_sync();
      const bool live = lane < qn;
      const uint32_t lr = stack[live ? lane : 0];
      qn = 0;
      rest(S, live, lr);
      wave_lds_sync();
    }
    }
  }
  if (!star_barrier(P)) return;
  // B2: backward filters over the survivors
  const int64_t nsurv = (int64_t)__hip_atomic_load(P.state + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int blnb = log_blocks_for(nsurv);
  // the init kernel cleared the filters; the next one clears what this run sets
  if (gtid < P.ndims) P.state[64 + gtid] = (unsigned long long)blnb;
  for (int d = 0; d < P.ndims; ++d)
    for (int64_t i = gtid; i < nsurv; i += gthreads)
      bf_insert(P.dim[d].bbf, blnb, s_masks, P.s_fk[(int64_t)d * P.fact_rows + i]);
  if (!star_barrier(P)) return;
  // C: kept dimension rows that pass the backward filter go into the hash tables
  for (int d = 0; d < P.ndims; ++d) {
    const StarDimDev& D = P.dim[d];
    const uint64_t hm = star_ht_mask(D, s_kept_rows[d]);
    for (int64_t r = gtid; r < D.rows; r += gthreads) {
      if (!D.kept[r]) continue;
      int64_t k = D.key[r];
      if (!bf_find(D.bbf, blnb, s_masks, k)) continue;
      ht_insert(D.ht, hm, k, (uint32_t)r, D.gattr ? (uint32_t)D.gattr[r] : 0u);
    }
  }
  if (!star_barrier(P)) return;
  // D: join the survivors with every dimension and aggregate
  __shared__ uint64_t hms[kStarMaxDims];
  if (threadIdx.x < P.ndims) hms[threadIdx.x] = star_ht_mask(P.dim[threadIdx.x], s_kept_rows[threadIdx.x]);
  __syncthreads();
  for (int64_t i = gtid; i < nsurv; i += gthreads) {
    uint32_t g[2] = {0, 0};
    bool ok = true;
    for (int j = 0; j < P.ndims && ok; ++j) {
      int d = P.order[j];
      uint32_t pay = 0;
      ok = ht_find(P.dim[d].ht, hms[d], P.s_fk[(int64_t)d * P.fact_rows + i], &pay);
      if (d == P.gdim[0]) g[0] = pay;
      if (d == P.gdim[1]) g[1] = pay;
    }
    if (ok) group_add(P, (uint64_t)g[0] | ((uint64_t)g[1] << 32), P.s_meas[i], P.s_idx[i]);
  }
}

__global__ void star_init_kernel(StarDev P) {
  const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t gthreads = (int64_t)gridDim.x * blockDim.x;
  if (gtid < 64) P.state[gtid] = 0;
  // the backward filters, as far as the run before may have set them (the
  // persistent kernel inserts into them without clearing first)
  for (int d = 0; d < P.ndims; ++d) {
    uint64_t* f = P.dim[d].bbf;
    const unsigned long long set = P.state[64 + d];
    const int64_t words = int64_t(1) << (set < (unsigned long long)P.bbf_cap_lnb ? (int)set : P.bbf_cap_lnb);
    for (int64_t b = gtid; b < words; b += gthreads) f[b] = 0;
  }
  // A1: filter the dimensions: kept flags, and kept rows per dimension per
  // workgroup (no device atomics: same-address atomics serialize)
  __shared__ unsigned long long s_cnt;
  for (int d = 0; d < P.ndims; ++d) {
    const StarDimDev& D = P.dim[d];
    unsigned long long kept = 0;
    for (int64_t r0 = gtid; r0 < D.rows; r0 += gthreads * kStarInitRows) {
      int32_t av[kStarInitRows];
#pragma unroll
      for (int j = 0; j < kStarInitRows; ++j) {
        const int64_t r = r0 + (int64_t)j * gthreads;
        av[j] = r < D.rows ? D.attr[r] : 0;
      }
#pragma unroll
      for (int j = 0; j < kStarInitRows; ++j) {
        const int64_t r = r0 + (int64_t)j * gthreads;
        if (r >= D.rows) continue;
        const bool k = star_cmp(av[j], D.op, D.cst);
        D.kept[r] = k ? 1 : 0;
        kept += k ? 1 : 0;
      }
    }
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    for (int off = 32; off > 0; off >>= 1) kept += __shfl_xor(kept, off);
    if ((threadIdx.x & 63) == 0 && kept) atomicAdd(&s_cnt, kept);
    __syncthreads();
    if (threadIdx.x == 0) P.kept_part[(int64_t)d * gridDim.x + blockIdx.x] = s_cnt;
    __syncthreads();   // s_cnt is reused for the next dimension
  }
  for (int64_t s = gtid; s <= (int64_t)P.g_mask + 1; s += gthreads) {
    P.g_keys[s] = kStarEmptyKey;
    P.g_cnt[s] = 0;
    P.g_sum[s] = 0;
    P.g_first[s] = LLONG_MAX;
  }
  // the forward filters, whole (a run uses a prefix sized by its kept rows)
  for (int d = 0; d < P.ndims; ++d) {
    uint64_t* f = P.dim[d].fbf;
    const int64_t words = int64_t(1) << P.dim[d].fbf_lnb;
    for (int64_t b = gtid; b < words; b += gthreads) f[b] = 0;
  }
}

// Ranks the groups by first appearance and writes the result columns.  One
// workgroup: the used slots are collected in LDS and sorted by (first row,
// slot); beyond kStarSortGroups groups each group counts the groups that came
// first instead.
constexpr int kStarSortGroups = 8192;
constexpr int kStarEmitScan = 8;   // group slots a thread per trip in the emit's scan

__device__ __forceinline__ void star_write_group(const StarDev& P, int64_t s, int64_t cap, int64_t rank,
                                                 int32_t* out_g0, int32_t* out_g1, int64_t* out_cnt,
                                                 int64_t* out_sum, int64_t max_groups) {
  if (rank >= max_groups) {   // more groups than the result buffers hold
    __hip_atomic_store(P.state + 32, 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  uint64_t k = (s == cap - 1) ? kStarEmptyKey : P.g_keys[s];
  out_g0[rank] = (int32_t)(uint32_t)(k & 0xFFFFFFFFULL);
  out_g1[rank] = (int32_t)(uint32_t)(k >> 32);
  out_cnt[rank] = (int64_t)P.g_cnt[s];
  out_sum[rank] = (int64_t)P.g_sum[s];
}

__global__ __launch_bounds__(1024) void star_emit_kernel(StarDev P, int32_t* out_g0, int32_t* out_g1,
                                                         int64_t* out_cnt, int64_t* out_sum,
                                                         unsigned long long* out_n, int64_t max_groups) {
  __shared__ unsigned long long skey[kStarSortGroups];
  __shared__ unsigned s_n;
  const int64_t cap = (int64_t)P.g_mask + 2;
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  // the used slots, kStarEmitScan a thread per trip with their loads issued
  // together; one LDS atomic per wave and trip item
  const int lane = threadIdx.x & 63;
  for (int64_t s0 = threadIdx.x; s0 < cap; s0 += (int64_t)blockDim.x * kStarEmitScan) {
    long long f[kStarEmitScan];
#pragma unroll
    for (int j = 0; j < kStarEmitScan; ++j) {
      const int64_t s = s0 + (int64_t)j * blockDim.x;
      f[j] = s < cap ? P.g_first[s] : LLONG_MAX;
    }
#pragma unroll
    for (int j = 0; j < kStarEmitScan; ++j) {
      const bool used = f[j] != LLONG_MAX;
      const uint64_t m = __ballot(used);
      if (!m) continue;
      unsigned base = 0;
      if (lane == __builtin_ctzll(m)) base = atomicAdd(&s_n, (unsigned)__popcll(m));
      base = __shfl(base, __builtin_ctzll(m));
      const unsigned i = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
      if (used && i < kStarSortGroups)
        skey[i] = ((unsigned long long)f[j] << 22) | (unsigned long long)(s0 + (int64_t)j * blockDim.x);
    }
  }
  __syncthreads();
  const unsigned n = s_n;
  if (n <= blockDim.x) {
    // one group a thread: its rank is the number of smaller keys (keys are
    // distinct: the slot is in the low bits); every thread reads the same
    // entry at a time, an LDS broadcast
    const unsigned long long mine = threadIdx.x < n ? skey[threadIdx.x] : 0ULL;
    unsigned rank = 0;
    for (unsigned j = 0; j < n; ++j) rank += skey[j] < mine ? 1u : 0u;
    if (threadIdx.x < n)
      star_write_group(P, (int64_t)(mine & ((1ULL << 22) - 1)), cap, rank, out_g0, out_g1, out_cnt, out_sum,
                       max_groups);
  } else if (n > kStarSortGroups) {
    for (int64_t s = threadIdx.x; s < cap; s += blockDim.x) {
      long long f = P.g_first[s];
      if (f == LLONG_MAX) continue;
      int64_t rank = 0;
      for (int64_t t = 0; t < cap; ++t) rank += (P.g_first[t] < f) ? 1 : 0;
      star_write_group(P, s, cap, rank, out_g0, out_g1, out_cnt, out_sum, max_groups);
    }
  } else {
    unsigned np = 1;
    while (np < n) np <<= 1;
    for (unsigned i = n + threadIdx.x; i < np; i += blockDim.x) skey[i] = ~0ULL;
    __syncthreads();
    for (unsigned k = 2; k <= np; k <<= 1) {
      for (unsigned j = k >> 1; j > 0; j >>= 1) {
        for (unsigned i = threadIdx.x; i < np; i += blockDim.x) {
          unsigned l = i ^ j;
          if (l > i) {
            unsigned long long a = skey[i], b = skey[l];
            if ((a > b) == ((i & k) == 0)) {
              skey[i] = b;
              skey[l] = a;
            }
          }
        }
        __syncthreads();
      }
    }
    for (unsigned i = threadIdx.x; i < n; i += blockDim.x)
      star_write_group(P, (int64_t)(skey[i] & ((1ULL << 22) - 1)), cap, i, out_g0, out_g1, out_cnt, out_sum,
                       max_groups);
  }
  if (threadIdx.x == 0) *out_n = n < (unsigned long long)max_groups ? n : (unsigned long long)max_groups;
}

}  // namespace fpdb

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

struct fpdb_star {
  fpdb::Context* ctx = nullptr;
  fpdb_star_query q;
  std::vector<int64_t> parts;          // partition offsets (num_partitions + 1)
  std::vector<fpdb::StarSeg> segs;     // the fact rows' segments (host copy of dev.segs)
  bool split = false;                  // partitions in buffers of their own (one segment each)
  fpdb_star_tuning tune;
  fpdb::StarDev dev;
  std::vector<void*> owned;
  int grid = 0;
  int resident_per_cu = 0;             // workgroups of the persistent kernel one CU holds
  void (*kernel)(fpdb::StarDev) = nullptr;   // the persistent kernel for the tuned items per thread
  // result (device)
  int32_t* out_g0 = nullptr;
  int32_t* out_g1 = nullptr;
  int64_t* out_cnt = nullptr;
  int64_t* out_sum = nullptr;
  unsigned long long* out_n = nullptr;
  int64_t max_groups = 0;
  // sizes of the last plain run (a batched run keeps them in the state words)
  int last_mode = 0;                       // 1 plain, 2 batched
  int64_t kept[fpdb::kStarMaxDims] = {0};
  int fwd_lnb[fpdb::kStarMaxDims] = {0};
  int bwd_lnb = 0;
  int bwd_cap_lnb = 0;                 // backward filter capacity (log2 blocks)
  unsigned long long bwd_mark[fpdb::kStarMaxDims] = {};   // source of mark_backward_filters' copy
  int64_t survivors = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
};

namespace fpdb {

// Partition p's fact foreign keys for dimension d, and its measures (in the
// partition's own buffers, or its rows of the contiguous table).
static const int64_t* fact_fk_of(const fpdb_star* s, int p, int d) {
  if (s->split) return s->segs[p].fk[d];
  return s->segs[0].fk[d] + s->parts[p];
}
static const int64_t* fact_measure_of(const fpdb_star* s, int p) {
  if (s->split) return s->segs[p].meas;
  return s->segs[0].meas + s->parts[p];
}

static fpdb_status star_alloc(fpdb_star* s, uint64_t bytes, void** p) {
  FPDB_TRY(ctx_alloc(s->ctx, bytes, p));
  s->owned.push_back(*p);
  return FPDB_OK;
}

static int host_log_blocks(int64_t n) { return fpdb_bloom_blocked_log_num_blocks(n); }

static fpdb_column col_i64(const int64_t* p, int64_t n) {
  fpdb_column c{};
  c.type = FPDB_TYPE_INT64;
  c.length = n;
  c.values = const_cast<int64_t*>(p);
  return c;
}
static fpdb_column col_i32(const int32_t* p, int64_t n) {
  fpdb_column c{};
  c.type = FPDB_TYPE_INT32;
  c.length = n;
  c.values = const_cast<int32_t*>(p);
  return c;
}

// Per-partition launches of the operator entry points (the drop-in path).
// Records every backward filter as set over its whole capacity, so that the
// next persistent run's init kernel clears all of it: after preparation (the
// allocations hold anything) and after a per-partition run (which uses the
// filters too).
static fpdb_status mark_backward_filters(fpdb_star* s) {
  for (int d = 0; d < fpdb::kStarMaxDims; ++d) s->bwd_mark[d] = (unsigned long long)s->bwd_cap_lnb;
  FPDB_HIP(hipMemcpyAsync(s->dev.state + 64, s->bwd_mark, sizeof(s->bwd_mark), hipMemcpyHostToDevice,
                          s->ctx->stream));
  return FPDB_OK;
}

static fpdb_status star_plain(fpdb_star* s) {
  Context* c = s->ctx;
  fpdb_context* ctx = reinterpret_cast<fpdb_context*>(c);
  const fpdb_star_query& q = s->q;
  const int nd = q.num_dims;
  const int64_t N = q.fact_rows;
  StarDev& P = s->dev;
  // dimensions: filter, forward filters, kept rows
  std::vector<int64_t> kept_n(nd);
  std::vector<int64_t*> kept_key(nd);
  std::vector<int32_t*> kept_g(nd);
  uint8_t* bm = nullptr;
  int64_t max_dim = 0;
  for (int d = 0; d < nd; ++d) max_dim = std::max(max_dim, q.dims[d].key.length);
  std::vector<void*> tmp;
  auto talloc = [&](uint64_t bytes, void** p) -> fpdb_status {
    FPDB_TRY(ctx_alloc(c, bytes, p));
    tmp.push_back(*p);
    return FPDB_OK;
  };
  auto cleanup = [&]() {
    for (void* p : tmp) ctx_free(c, p);
    tmp.clear();
  };
  fpdb_status st = FPDB_OK;
#define STAR_TRY(x)               \
  do {                            \
    st = (x);                     \
    if (st != FPDB_OK) {          \
      cleanup();                  \
      return st;                  \
    }                             \
  } while (0)
  STAR_TRY(talloc((uint64_t)((std::max<int64_t>(max_dim, N) + 63) / 64) * 8, reinterpret_cast<void**>(&bm)));
  int64_t* d_cnt = nullptr;
  STAR_TRY(talloc(64, reinterpret_cast<void**>(&d_cnt)));
  for (int d = 0; d < nd; ++d) {
    const fpdb_star_dimension& D = q.dims[d];
    int64_t rows = D.key.length;
    STAR_TRY(fpdb_compare_bitmap(ctx, &D.filter_attr, D.filter_op, D.filter_constant, 0.0, nullptr, rows, bm, d_cnt));
    int64_t k = 0;
    FPDB_HIP(hipMemcpyAsync(&k, d_cnt, 8, hipMemcpyDeviceToHost, c->stream));
    FPDB_HIP(hipStreamSynchronize(c->stream));
    kept_n[d] = k;
    s->kept[d] = k;
    s->fwd_lnb[d] = host_log_blocks(k);
    STAR_TRY(fpdb_bloom_blocked_build(ctx, &D.key, 1, rows, bm, s->fwd_lnb[d], P.dim[d].fbf));
    STAR_TRY(talloc((uint64_t)std::max<int64_t>(rows, 1) * 8, reinterpret_cast<void**>(&kept_key[d])));
    STAR_TRY(talloc((uint64_t)std::max<int64_t>(rows, 1) * 4, reinterpret_cast<void**>(&kept_g[d])));
    fpdb_column in[2] = {D.key, D.has_group_attr ? D.group_attr : D.filter_attr};
    fpdb_column out[2] = {col_i64(kept_key[d], rows), col_i32(kept_g[d], rows)};
    int64_t kk = 0;
    STAR_TRY(fpdb_filter_compact(ctx, in, 2, rows, bm, out, &kk));
  }
  // forward pass, one partition at a time, filters in order
  const int np = (int)s->parts.size() - 1;
  std::vector<std::vector<int64_t*>> surv(np);   // per partition: nd fk columns + measure
  std::vector<int64_t> surv_n(np, 0);
  int64_t total = 0;
  for (int p = 0; p < np; ++p) {
    int64_t lo = s->parts[p], n = s->parts[p + 1] - lo;
    std::vector<fpdb_column> cur;
    for (int d = 0; d < nd; ++d) cur.push_back(col_i64(fact_fk_of(s, p, d), n));
    cur.push_back(col_i64(fact_measure_of(s, p), n));
    for (int i = 0; i < nd && n > 0; ++i) {
      int d = q.filter_order[i];
      std::vector<int64_t*> outp(nd + 1);
      std::vector<fpdb_column> out;
      for (int j = 0; j <= nd; ++j) {
        STAR_TRY(talloc((uint64_t)std::max<int64_t>(n, 1) * 8, reinterpret_cast<void**>(&outp[j])));
        out.push_back(col_i64(outp[j], n));
      }
      int key = d;
      int64_t kept = 0;
      STAR_TRY(fpdb_bloom_blocked_filter(ctx, P.dim[d].fbf, s->fwd_lnb[d], cur.data(), nd + 1, &key, 1, n, nullptr,
                                         out.data(), &kept));
      n = kept;
      cur.clear();
      for (int j = 0; j <= nd; ++j) cur.push_back(col_i64(outp[j], n));
      surv[p] = outp;
    }
    if (surv[p].empty()) {
      // zero-row partition or no filters: keep the (empty) input columns
      for (int j = 0; j <= nd; ++j) surv[p].push_back(static_cast<int64_t*>(cur[j].values));
    }
    surv_n[p] = n;
    total += n;
  }
  // backward filters over the union of the survivors
  s->bwd_lnb = host_log_blocks(total);
  s->survivors = total;
  for (int d = 0; d < nd; ++d) {
    STAR_TRY(launch_fill(c, P.dim[d].bbf, 0, (uint64_t)8 << s->bwd_lnb));
    STAR_TRY(mark_backward_filters(s));   // for the next persistent run
    for (int p = 0; p < np; ++p) {
      if (!surv_n[p]) continue;
      fpdb_column k = col_i64(surv[p][d], surv_n[p]);
      STAR_TRY(fpdb_bloom_blocked_insert(ctx, &k, 1, surv_n[p], nullptr, s->bwd_lnb, P.dim[d].bbf));
    }
  }
  // reduce the dimensions with the backward filters, build the hash tables
  std::vector<fpdb_hash_table*> tables(nd, nullptr);
  auto drop_tables = [&]() {
    for (auto t : tables)
      if (t) fpdb_hj_table_destroy(t);
  };
  for (int d = 0; d < nd; ++d) {
    int64_t rows = kept_n[d];
    int64_t* rk = nullptr;
    int32_t* rg = nullptr;
    STAR_TRY(talloc((uint64_t)std::max<int64_t>(rows, 1) * 8, reinterpret_cast<void**>(&rk)));
    STAR_TRY(talloc((uint64_t)std::max<int64_t>(rows, 1) * 4, reinterpret_cast<void**>(&rg)));
    fpdb_column in[2] = {col_i64(kept_key[d], rows), col_i32(kept_g[d], rows)};
    fpdb_column out[2] = {col_i64(rk, rows), col_i32(rg, rows)};
    int key = 0;
    int64_t kept = 0;
    st = fpdb_bloom_blocked_filter(ctx, P.dim[d].bbf, s->bwd_lnb, in, 2, &key, 1, rows, nullptr, out, &kept);
    if (st == FPDB_OK) {
      fpdb_column kc = col_i64(rk, kept), pc = col_i32(rg, kept);
      st = fpdb_hj_build(ctx, &kc, 1, kept, &pc, 0, &tables[d]);
    }
    if (st != FPDB_OK) {
      drop_tables();
      cleanup();
      return st;
    }
  }
  // joins and group-by, one partition at a time
  int32_t ktypes[2] = {FPDB_TYPE_INT32, FPDB_TYPE_INT32};
  fpdb_agg_spec aggs[2] = {{FPDB_AGG_COUNT_ALL, FPDB_TYPE_INT64, 0, 0}, {FPDB_AGG_SUM, FPDB_TYPE_INT64, 0, 0}};
  fpdb_group_table* gt = nullptr;
  st = fpdb_groupby_create(ctx, ktypes, 2, aggs, 2, s->max_groups, &gt);
  if (st != FPDB_OK) {
    drop_tables();
    cleanup();
    return st;
  }
  for (int p = 0; p < np && st == FPDB_OK; ++p) {
    int64_t n = surv_n[p];
    if (!n) continue;
    // working columns: the fact columns still needed, then gathered group attributes
    std::vector<fpdb_column> cur;
    std::vector<int> what;   // >= 0: fk of dim, -1: measure, -2 - g: group attribute g
    for (int d = 0; d < nd; ++d) {
      cur.push_back(col_i64(surv[p][d], n));
      what.push_back(d);
    }
    cur.push_back(col_i64(surv[p][nd], n));
    what.push_back(-1);
    for (int j = 0; j < nd && st == FPDB_OK && n > 0; ++j) {
      int d = q.filter_order[j];
      int keyi = 0;
      while (what[keyi] != d) ++keyi;
      std::vector<fpdb_column> probe_out;
      std::vector<int> nwhat;
      for (size_t i = 0; i < cur.size(); ++i) {
        if ((int)i == keyi) continue;
        probe_out.push_back(cur[i]);
        nwhat.push_back(what[i]);
      }
      std::vector<fpdb_column> build_out;
      int g = (d == q.group_dims[0]) ? 0 : (d == q.group_dims[1]) ? 1 : -1;
      if (g >= 0) {
        fpdb_column pc{};
        pc.type = FPDB_TYPE_INT32;
        build_out.push_back(pc);   // inline payload = group attribute
        nwhat.push_back(-2 - g);
      }
      std::vector<fpdb_column> out;
      for (size_t i = 0; i < probe_out.size() + build_out.size(); ++i) {
        void* pbuf = nullptr;
        st = talloc((uint64_t)n * 8, &pbuf);
        if (st != FPDB_OK) break;
        fpdb_column oc = i < probe_out.size() ? probe_out[i] : build_out[i - probe_out.size()];
        oc.values = pbuf;
        oc.length = n;
        oc.offset = 0;
        oc.validity = nullptr;
        out.push_back(oc);
      }
      if (st != FPDB_OK) break;
      int64_t produced = 0;
      st = fpdb_hj_probe(ctx, tables[d], FPDB_JOIN_INNER, &cur[keyi], 1, n, probe_out.data(), (int)probe_out.size(),
                         build_out.data(), (int)build_out.size(), out.data(), n, &produced);
      n = produced;
      for (auto& oc : out) oc.length = n;
      cur = out;
      what = nwhat;
    }
    if (st != FPDB_OK || n == 0) continue;
    fpdb_column keys[2], vals[2];
    for (size_t i = 0; i < cur.size(); ++i) {
      if (what[i] == -2) keys[0] = cur[i];
      if (what[i] == -3) keys[1] = cur[i];
      if (what[i] == -1) vals[1] = cur[i];
    }
    vals[0] = fpdb_column{};
    st = fpdb_groupby_consume(ctx, gt, keys, vals, n);
  }
  if (st == FPDB_OK) {
    int64_t ng = 0;
    st = fpdb_groupby_num_groups(ctx, gt, &ng);
    if (st == FPDB_OK && ng > s->max_groups) st = fail(FPDB_ERR_CAPACITY, "star join: too many groups");
    if (st == FPDB_OK) {
      fpdb_column ok[2] = {col_i32(s->out_g0, ng), col_i32(s->out_g1, ng)};
      int64_t* cnt_tmp = s->out_cnt;
      fpdb_column oa[2] = {col_i64(cnt_tmp, ng), col_i64(s->out_sum, ng)};
      int64_t written = 0;
      st = fpdb_groupby_finalize(ctx, gt, ok, oa, &written);
      unsigned long long hn = (unsigned long long)written;
      if (st == FPDB_OK) {
        hipError_t e = hipMemcpyAsync(s->out_n, &hn, 8, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) st = fail(FPDB_ERR_HIP, hipGetErrorString(e));
      }
    }
  }
  fpdb_groupby_destroy(gt);
  drop_tables();
  cleanup();
#undef STAR_TRY
  return st;
}

static fpdb_status star_batched_enqueue(fpdb_star* s) {
  Context* c = s->ctx;
  // every workgroup of the persistent kernel must be resident at once for its
  // grid barriers: the grid never exceeds what the CUs hold at its LDS and
  // register use
  FPDB_CHECK(s->grid >= 1 && s->grid <= s->resident_per_cu * std::max(1, c->num_cus), FPDB_ERR_INTERNAL,
             "star join: persistent grid exceeds the resident workgroups");
  hipLaunchKernelGGL(star_init_kernel, dim3(s->dev.init_blocks), dim3(kStarInitThreads), 0, c->stream, s->dev);
  FPDB_HIP(hipGetLastError());
  {
    ProfScope ps(c, "star_persistent");
    hipLaunchKernelGGL(s->kernel, dim3(s->grid), dim3(kStarThreads), kStarDynamicLds, c->stream, s->dev);
  }
  FPDB_HIP(hipGetLastError());
  hipLaunchKernelGGL(star_emit_kernel, dim3(1), dim3(1024), 0, c->stream, s->dev, s->out_g0, s->out_g1, s->out_cnt,
                     s->out_sum, s->out_n, s->max_groups);
  FPDB_HIP(hipGetLastError());
  return FPDB_OK;
}

}  // namespace fpdb

using namespace fpdb;

extern "C" {

fpdb_status fpdb_star_prepare(fpdb_context* ctx, const fpdb_star_query* q, const fpdb_star_tuning* tuning,
                              fpdb_star** out) {
  FPDB_CHECK(ctx && q && out, FPDB_ERR_INVALID, "NULL argument");
  *out = nullptr;
  FPDB_CHECK(q->num_dims >= 1 && q->num_dims <= kStarMaxDims, FPDB_ERR_INVALID, "star join: 1..8 dimensions");
  FPDB_CHECK(q->num_partitions >= 1 && q->partition_offsets, FPDB_ERR_INVALID, "star join: partitions needed");
  FPDB_CHECK(q->fact_rows >= 0 && q->fact_rows < (int64_t(1) << 32), FPDB_ERR_INVALID,
             "star join: at most 2^32 - 1 fact rows (rows within a segment are 32-bit)");
  FPDB_CHECK(q->partition_offsets[0] == 0 && q->partition_offsets[q->num_partitions] == q->fact_rows,
             FPDB_ERR_INVALID, "star join: partition offsets must cover the fact table");
  for (int p = 0; p < q->num_partitions; ++p)
    FPDB_CHECK(q->partition_offsets[p] <= q->partition_offsets[p + 1], FPDB_ERR_INVALID,
               "star join: partition offsets must not decrease");
  const bool split = q->partition_fact_fk != nullptr;
  if (split) {
    FPDB_CHECK(q->partition_measure, FPDB_ERR_INVALID, "star join: partition_fact_fk without partition_measure");
    for (int p = 0; p < q->num_partitions; ++p) {
      const int64_t rows = q->partition_offsets[p + 1] - q->partition_offsets[p];
      const fpdb_column& m = q->partition_measure[p];
      FPDB_CHECK(m.type == FPDB_TYPE_INT64 && m.length >= rows && (m.values || rows == 0) && !m.validity,
                 FPDB_ERR_INVALID, "star join: non-null int64 partition measure columns needed");
      for (int d = 0; d < q->num_dims; ++d) {
        const fpdb_column& f = q->partition_fact_fk[(size_t)p * q->num_dims + d];
        FPDB_CHECK(f.type == FPDB_TYPE_INT64 && f.length >= rows && (f.values || rows == 0) && !f.validity,
                   FPDB_ERR_INVALID, "star join: non-null int64 partition foreign-key columns needed");
      }
    }
  } else {
    FPDB_CHECK(q->fact_measure.type == FPDB_TYPE_INT64 && q->fact_measure.length >= q->fact_rows &&
                   q->fact_measure.values,
               FPDB_ERR_INVALID, "star join: int64 measure column needed");
  }
  FPDB_CHECK(q->group_dims[0] >= 0 && q->group_dims[0] < q->num_dims && q->group_dims[1] >= 0 &&
                 q->group_dims[1] < q->num_dims && q->group_dims[0] != q->group_dims[1],
             FPDB_ERR_INVALID, "star join: two distinct group dimensions");
  std::vector<int> seen(q->num_dims, 0);
  for (int i = 0; i < q->num_dims; ++i) {
    int d = q->filter_order[i];
    FPDB_CHECK(d >= 0 && d < q->num_dims && !seen[d], FPDB_ERR_INVALID, "star join: filter_order is not a permutation");
    seen[d] = 1;
    const fpdb_star_dimension& D = q->dims[d];
    FPDB_CHECK(D.key.type == FPDB_TYPE_INT64 && D.key.values && D.key.offset == 0 && !D.key.validity, FPDB_ERR_INVALID,
               "star join: dimension keys are non-null int64 columns");
    FPDB_CHECK(D.filter_attr.type == FPDB_TYPE_INT32 && D.filter_attr.length >= D.key.length && D.filter_attr.values &&
                   D.filter_attr.offset == 0,
               FPDB_ERR_INVALID, "star join: int32 filter attribute needed");
    FPDB_CHECK(D.filter_op >= FPDB_CMP_LT && D.filter_op <= FPDB_CMP_NE, FPDB_ERR_INVALID, "star join: bad filter op");
    FPDB_CHECK(split || (D.fact_fk.type == FPDB_TYPE_INT64 && D.fact_fk.length >= q->fact_rows &&
                         D.fact_fk.values && D.fact_fk.offset == 0 && !D.fact_fk.validity),
               FPDB_ERR_INVALID, "star join: non-null int64 fact foreign keys needed");
    FPDB_CHECK(D.key.length < (int64_t)kStarEmptyRow, FPDB_ERR_INVALID, "star join: dimension too large");
  }
  for (int g = 0; g < 2; ++g) {
    const fpdb_star_dimension& D = q->dims[q->group_dims[g]];
    FPDB_CHECK(D.has_group_attr && D.group_attr.type == FPDB_TYPE_INT32 && D.group_attr.values &&
                   D.group_attr.length >= D.key.length && D.group_attr.offset == 0,
               FPDB_ERR_INVALID, "star join: group dimensions need an int32 group attribute");
  }
  FPDB_CHECK(q->max_groups >= 1 && q->max_groups <= (1 << 20), FPDB_ERR_INVALID, "star join: max_groups in [1, 2^20]");
  Context* c = as_ctx(ctx);
  auto* s = new fpdb_star();
  s->ctx = c;
  s->q = *q;
  s->parts.assign(q->partition_offsets, q->partition_offsets + q->num_partitions + 1);
  s->q.partition_offsets = nullptr;
  // the fact rows' segments: the contiguous table, or one per partition
  s->split = split;
  if (split) {
    for (int p = 0; p < q->num_partitions; ++p) {
      StarSeg g{};
      g.base = s->parts[p];
      g.rows = s->parts[p + 1] - s->parts[p];
      const fpdb_column& m = q->partition_measure[p];
      g.meas = static_cast<const int64_t*>(m.values) + m.offset;
      for (int d = 0; d < q->num_dims; ++d) {
        const fpdb_column& f = q->partition_fact_fk[(size_t)p * q->num_dims + d];
        g.fk[d] = static_cast<const int64_t*>(f.values) + f.offset;
      }
      s->segs.push_back(g);
    }
  } else {
    StarSeg g{};
    g.base = 0;
    g.rows = q->fact_rows;
    g.meas = static_cast<const int64_t*>(q->fact_measure.values);
    for (int d = 0; d < q->num_dims; ++d) g.fk[d] = static_cast<const int64_t*>(q->dims[d].fact_fk.values);
    s->segs.push_back(g);
  }
  s->q.partition_fact_fk = nullptr;   // caller memory: only the segments are kept
  s->q.partition_measure = nullptr;
  s->tune = tuning ? *tuning : fpdb_star_tuning{0, 0, 1, 0};
  if (s->tune.ht_load_shift < 1 || s->tune.ht_load_shift > 3) s->tune.ht_load_shift = 1;
  FPDB_CHECK(s->tune.items_per_thread == 0 || s->tune.items_per_thread == 4 || s->tune.items_per_thread == 8 ||
                 s->tune.items_per_thread == 16,
             FPDB_ERR_INVALID, "star join: items_per_thread must be 0, 4, 8 or 16");
  FPDB_CHECK(s->tune.barrier_timeout_us >= 0, FPDB_ERR_INVALID, "star join: negative barrier timeout");
  s->max_groups = q->max_groups;
  StarDev& P = s->dev;
  std::memset(&P, 0, sizeof(P));
  P.ndims = q->num_dims;
  for (int i = 0; i < q->num_dims; ++i) P.order[i] = q->filter_order[i];
  P.gdim[0] = q->group_dims[0];
  P.gdim[1] = q->group_dims[1];
  P.fact_rows = q->fact_rows;
  P.nsegs = (int32_t)s->segs.size();
  fpdb_status st = FPDB_OK;
  int blnb = host_log_blocks(q->fact_rows);
  for (int d = 0; d < q->num_dims && st == FPDB_OK; ++d) {
    const fpdb_star_dimension& D = q->dims[d];
    StarDimDev& X = P.dim[d];
    X.key = static_cast<const int64_t*>(D.key.values);
    X.attr = static_cast<const int32_t*>(D.filter_attr.values);
    X.gattr = D.has_group_attr ? static_cast<const int32_t*>(D.group_attr.values) : nullptr;
    X.rows = D.key.length;
    X.cst = D.filter_constant;
    X.op = D.filter_op;
    int64_t cap = 64;
    while (cap < (X.rows << s->tune.ht_load_shift)) cap <<= 1;
    X.ht_mask = (uint64_t)cap - 1;
    X.ht_shift = s->tune.ht_load_shift;
    st = star_alloc(s, (uint64_t)std::max<int64_t>(X.rows, 1), reinterpret_cast<void**>(&X.kept));
    X.fbf_lnb = host_log_blocks(X.rows);
    if (st == FPDB_OK) st = star_alloc(s, (uint64_t)8 << X.fbf_lnb, reinterpret_cast<void**>(&X.fbf));
    if (st == FPDB_OK) st = star_alloc(s, (uint64_t)8 << blnb, reinterpret_cast<void**>(&X.bbf));
    if (st == FPDB_OK) st = star_alloc(s, (uint64_t)cap * sizeof(StarSlot), reinterpret_cast<void**>(&X.ht));
  }
  int64_t N = std::max<int64_t>(q->fact_rows, 1);
  if (st == FPDB_OK) st = star_alloc(s, (uint64_t)N * 8, reinterpret_cast<void**>(&P.s_idx));
  if (st == FPDB_OK) st = star_alloc(s, (uint64_t)N * 8 * q->num_dims, reinterpret_cast<void**>(&P.s_fk));
  if (st == FPDB_OK) st = star_alloc(s, (uint64_t)N * 8, reinterpret_cast<void**>(&P.s_meas));
  if (st == FPDB_OK) st = star_alloc(s, kStarStateWords * 8, reinterpret_cast<void**>(&P.state));
  if (st == FPDB_OK) st = star_alloc(s, s->segs.size() * sizeof(StarSeg), reinterpret_cast<void**>(const_cast<StarSeg**>(&P.segs)));
  if (st == FPDB_OK) st = hipMemcpy(const_cast<StarSeg*>(P.segs), s->segs.data(), s->segs.size() * sizeof(StarSeg),
                                    hipMemcpyHostToDevice) == hipSuccess ? FPDB_OK : fail(FPDB_ERR_HIP, "star join: segment copy");
  int64_t gcap = 64;
  while (gcap < 2 * q->max_groups) gcap <<= 1;
  P.g_mask = (uint64_t)gcap - 1;
  if (st == FPDB_OK) st = star_alloc(s, (uint64_t)(gcap + 1) * 8, reinterpret_cast<void**>(&P.g_keys));
  if (st == FPDB_OK) st = star_alloc(s, (uint64_t)(gcap + 1) * 8, reinterpret_cast<void**>(&P.g_cnt));
  if (st == FPDB_OK) st = star_alloc(s, (uint64_t)(gcap + 1) * 8, reinterpret_cast<void**>(&P.g_sum));
  if (st == FPDB_OK) st = star_alloc(s, (uint64_t)(gcap + 1) * 8, reinterpret_cast<void**>(&P.g_first));
  if (st == FPDB_OK) st = star_alloc(s, (uint64_t)q->max_groups * 4, reinterpret_cast<void**>(&s->out_g0));
  if (st == FPDB_OK) st = star_alloc(s, (uint64_t)q->max_groups * 4, reinterpret_cast<void**>(&s->out_g1));
  if (st == FPDB_OK) st = star_alloc(s, (uint64_t)q->max_groups * 8, reinterpret_cast<void**>(&s->out_cnt));
  if (st == FPDB_OK) st = star_alloc(s, (uint64_t)q->max_groups * 8, reinterpret_cast<void**>(&s->out_sum));
  if (st == FPDB_OK) st = star_alloc(s, 8, reinterpret_cast<void**>(&s->out_n));
  if (st == FPDB_OK) st = launch_fill(c, P.state, 0, 64 * 8);   // no error recorded before the first run
  s->bwd_cap_lnb = blnb;
  P.bbf_cap_lnb = blnb;
  {
    // init workgroups: kStarInitRows rows a thread of the largest dimension
    int64_t max_rows = 0;
    for (int d = 0; d < q->num_dims; ++d) max_rows = std::max<int64_t>(max_rows, P.dim[d].rows);
    const int64_t per = (int64_t)kStarInitThreads * kStarInitRows;
    P.init_blocks = (int32_t)std::min<int64_t>(std::max<int64_t>((max_rows + per - 1) / per, 256), kStarInitMaxBlocks);
  }
  if (st == FPDB_OK)
    st = star_alloc(s, (uint64_t)kStarMaxDims * P.init_blocks * 8, reinterpret_cast<void**>(&P.kept_part));
  if (st == FPDB_OK) st = mark_backward_filters(s);   // the allocations hold anything
  if (st != FPDB_OK) {
    fpdb_star_destroy(s);
    return st;
  }
  // persistent grid: every workgroup resident at once
  switch (s->tune.items_per_thread) {
    case 4: s->kernel = star_persistent_kernel<4>; break;
    case 16: s->kernel = star_persistent_kernel<16>; break;
    default: s->kernel = star_persistent_kernel<kStarItems>; break;
  }
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, s->kernel, kStarThreads, kStarDynamicLds);
  if (e != hipSuccess || per_cu < 1) {
    fpdb_star_destroy(s);
    return fail(FPDB_ERR_HIP, "star join: occupancy query failed");
  }
  int want = s->tune.blocks_per_cu > 0 ? s->tune.blocks_per_cu : 4;
  s->resident_per_cu = per_cu;
  s->grid = std::max(1, c->num_cus) * std::min(want, per_cu);
  // grid barrier wait bound in ticks of the device's constant-rate wall clock
  int clock_khz = 0;
  e = hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, c->device);
  if (e != hipSuccess || clock_khz <= 0) {
    fpdb_star_destroy(s);
    return fail(FPDB_ERR_HIP, "star join: wall clock rate query failed");
  }
  const uint64_t us = s->tune.barrier_timeout_us > 0 ? (uint64_t)s->tune.barrier_timeout_us : kStarBarrierTimeoutUs;
  s->dev.barrier_ticks = us * (uint64_t)clock_khz / 1000;
  *out = s;
  return FPDB_OK;
}

fpdb_status fpdb_star_destroy(fpdb_star* s) {
  if (!s) return FPDB_OK;
  if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  for (void* p : s->owned) ctx_free(s->ctx, p);
  delete s;
  return FPDB_OK;
}

fpdb_status fpdb_star_run_plain(fpdb_context* ctx, fpdb_star* s) {
  FPDB_CHECK(ctx && s && as_ctx(ctx) == s->ctx, FPDB_ERR_INVALID, "star join: bad context");
  s->last_mode = 1;
  return star_plain(s);
}

fpdb_status fpdb_star_run_batched(fpdb_context* ctx, fpdb_star* s) {
  FPDB_CHECK(ctx && s && as_ctx(ctx) == s->ctx, FPDB_ERR_INVALID, "star join: bad context");
  s->last_mode = 2;
  return star_batched_enqueue(s);
}

fpdb_status fpdb_star_graph_capture(fpdb_context* ctx, fpdb_star* s) {
  FPDB_CHECK(ctx && s && as_ctx(ctx) == s->ctx, FPDB_ERR_INVALID, "star join: bad context");
  Context* c = s->ctx;
  FPDB_CHECK(c->own_stream, FPDB_ERR_INVALID, "star join: graph capture needs the context's own stream");
  if (s->graph_exec) return FPDB_OK;
  bool prof = c->profiling;
  c->profiling = false;
  FPDB_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  fpdb_status st = star_batched_enqueue(s);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  c->profiling = prof;
  FPDB_TRY(st);
  if (e != hipSuccess) return fail(FPDB_ERR_HIP, std::string("graph capture failed: ") + hipGetErrorString(e));
  s->graph = g;
  FPDB_HIP(hipGraphInstantiate(&s->graph_exec, g, nullptr, nullptr, 0));
  return FPDB_OK;
}

fpdb_status fpdb_star_graph_launch(fpdb_context* ctx, fpdb_star* s) {
  FPDB_CHECK(ctx && s && s->graph_exec, FPDB_ERR_INVALID, "star join: no captured graph");
  ProfScope ps(s->ctx, "star_graph");
  s->last_mode = 2;
  FPDB_HIP(hipGraphLaunch(s->graph_exec, s->ctx->stream));
  return FPDB_OK;
}

fpdb_status fpdb_star_result(fpdb_context* ctx, fpdb_star* s, int32_t* g0, int32_t* g1, int64_t* count,
                             int64_t* sum, int64_t capacity, int64_t* num_groups) {
  FPDB_CHECK(ctx && s && num_groups, FPDB_ERR_INVALID, "NULL argument");
  Context* c = s->ctx;
  unsigned long long n = 0;
  unsigned long long err = 0;
  FPDB_HIP(hipMemcpyAsync(&n, s->out_n, 8, hipMemcpyDeviceToHost, c->stream));
  // the error word belongs to the batched form; the plain form reports errors from its own calls
  if (s->last_mode == 2) FPDB_HIP(hipMemcpyAsync(&err, s->dev.state + 32, 8, hipMemcpyDeviceToHost, c->stream));
  FPDB_HIP(hipStreamSynchronize(c->stream));
  FPDB_CHECK(err == 0, FPDB_ERR_INTERNAL, "star join: a grid barrier timed out or the group table overflowed");
  *num_groups = (int64_t)n;
  FPDB_CHECK((int64_t)n <= capacity, FPDB_ERR_CAPACITY, "star join: result capacity too small");
  if (n) {
    if (g0) FPDB_HIP(hipMemcpy(g0, s->out_g0, n * 4, hipMemcpyDeviceToHost));
    if (g1) FPDB_HIP(hipMemcpy(g1, s->out_g1, n * 4, hipMemcpyDeviceToHost));
    if (count) FPDB_HIP(hipMemcpy(count, s->out_cnt, n * 8, hipMemcpyDeviceToHost));
    if (sum) FPDB_HIP(hipMemcpy(sum, s->out_sum, n * 8, hipMemcpyDeviceToHost));
  }
  return FPDB_OK;
}

fpdb_status fpdb_star_filter(fpdb_context* ctx, fpdb_star* s, int dim, int backward, uint64_t* blocks,
                             int64_t capacity, int* log_num_blocks, int64_t* filter_input_rows) {
  FPDB_CHECK(ctx && s && log_num_blocks, FPDB_ERR_INVALID, "NULL argument");
  FPDB_CHECK(dim >= 0 && dim < s->q.num_dims, FPDB_ERR_INVALID, "star join: bad dimension");
  Context* c = s->ctx;
  FPDB_CHECK(s->last_mode != 0, FPDB_ERR_INVALID, "star join: not run yet");
  int64_t rows;
  if (s->last_mode == 2) {
    unsigned long long st[16];
    FPDB_HIP(hipMemcpyAsync(st, s->dev.state, sizeof(st), hipMemcpyDeviceToHost, c->stream));
    FPDB_HIP(hipStreamSynchronize(c->stream));
    rows = backward ? (int64_t)st[0] : (int64_t)st[1 + dim];
  } else {
    rows = backward ? s->survivors : s->kept[dim];
  }
  int lnb = fpdb_bloom_blocked_log_num_blocks(rows);
  *log_num_blocks = lnb;
  if (filter_input_rows) *filter_input_rows = rows;
  FPDB_CHECK(capacity >= (int64_t(1) << lnb) || !blocks, FPDB_ERR_CAPACITY, "star join: filter buffer too small");
  if (blocks)
    FPDB_HIP(hipMemcpy(blocks, backward ? s->dev.dim[dim].bbf : s->dev.dim[dim].fbf, (size_t)8 << lnb,
                       hipMemcpyDeviceToHost));
  return FPDB_OK;
}

}  // extern "C"
