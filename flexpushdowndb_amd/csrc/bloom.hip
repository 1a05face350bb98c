// Bloom filters of FlexPushdownDB's predicate transfer on gfx950:
//   * key hashing (Arrow 6 HashMultiColumn on fixed-width keys),
//   * the blocked filter (one 64-bit block per key, 57-bit masks),
//   * the legacy universal-hash filter,
//   * order-preserving filter compaction (count pass, tile scan, copy pass),
//     used plain (bitmap in) and after the bloom probe.
//
// Every kernel here streams its key columns once, coalesced: one wavefront
// covers 64 consecutive rows and the per-row test result becomes one 64-bit
// bitmap word by a wave ballot, which is exactly the LSB-first layout the
// reference's FindImp writes (fpdb-tuple/src/arrow/exec/BloomFilter.cpp:138-170).
#include "fpdb_internal.hpp"
#include "compact.hpp"

#include <algorithm>
#include <type_traits>
#include <random>
#include <vector>

namespace fpdb {

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int n) {
  return (x << n) | (x >> ((-n) & 63));
}

// BlockedBloomFilter::mask / block_id (BloomFilter.h:190-212) on a 32-bit hash.
__device__ __forceinline__ uint64_t bb_mask(const uint64_t* s_masks, uint32_t h) {
  return rotl64(s_masks[h & (kMaskNum - 1)], (h >> kMaskLogNum) & 63);
}
__device__ __forceinline__ uint64_t bb_block(uint32_t h, uint64_t nb_mask) {
  return ((uint64_t)h >> (kMaskLogNum + 6)) & nb_mask;
}

__device__ __forceinline__ void load_masks_lds(uint64_t* s_masks) {
  for (int i = threadIdx.x; i < kMaskNum; i += blockDim.x) s_masks[i] = c_masks.m[i];
  __syncthreads();
}

// ---------------------------------------------------------------------------
// hashing
// ---------------------------------------------------------------------------
__global__ void hash_keys_kernel(ColSet keys, int64_t n, uint32_t* __restrict__ out) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride)
    out[r] = hash_row(keys, r);
}

// ---------------------------------------------------------------------------
// blocked filter: insert
// ---------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(256) void blocked_insert_kernel(ColSet keys, int64_t n,
                                                             const uint8_t* __restrict__ select,
                                                             uint64_t* __restrict__ blocks,
                                                             uint64_t nb_mask) {
  __shared__ uint64_t s_masks[kMaskNum];
  load_masks_lds(s_masks);
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
    if (select && !bitmap_bit(select, r)) continue;
    uint32_t h = key_hash<MODE>(keys, r);
    __hip_atomic_fetch_or(&blocks[bb_block(h, nb_mask)], bb_mask(s_masks, h), __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT);
  }
}

// ---------------------------------------------------------------------------
// blocked filter: find with the filter in LDS
//
// A key's block index is bits 16..31 of its hash, so at most 2^16 blocks
// (512 KiB) are ever read.  Random 8-byte reads from the L2 run at about
// 220 G requests/s on this part, a quarter of what the key stream needs, so
// block reads go to LDS wherever the filter fits:
//   * up to 2^14 reachable blocks (128 KiB): every workgroup holds the whole
//     filter and tests its own range of keys;
//   * larger filters are cut into slices of 2^14 blocks.  Workgroups form
//     groups of G that read the same range of keys; member q holds slice q
//     and tests the keys whose block lies in it, plus -- when the G slices do
//     not cover the filter -- the keys of the remaining blocks b with
//     b % G == q, read from L2.  Each member writes partial bitmap words (bit i
//     set iff member q tested row i and it passed); a combine kernel ORs the G
//     partial words.  Workgroup b runs on XCD b % 8 (round-robin dispatch)
//     and the members of a group are b, b + 8, ..., b + 8(G-1): they share one
//     XCD's L2, so a key line one member fetched from HBM is an L2 hit for the
//     others.
// Every member loads and hashes every key of its range, so G trades VALU and
// texture-address work (G loads and hashes per key) against L2 gathers (keys
// of blocks outside the slices; a 64-lane gather of distinct lines costs the
// texture-address unit about 4x a coalesced load).  Measured on a 2^16-block
// filter with 60M keys: G = 1, 2 and 4 all take 0.28-0.30 ms, G = 1 bound by
// the gathers, G = 4 by VALU (33 instructions per 64 keys per member); G = 4
// is kept since it makes no L2 gathers at all.
constexpr int kSliceLog = 14;
constexpr int kSliceThreads = 1024;
#ifndef FPDB_SLICE_UNROLL
#define FPDB_SLICE_UNROLL 8   // bitmap words per wave per step (loads in flight together)
#endif
constexpr int kSliceUnroll = FPDB_SLICE_UNROLL;
#ifndef FPDB_FIND_GROUP
#define FPDB_FIND_GROUP 4     // most workgroups sharing one range of keys
#endif

inline int reachable_log_blocks(int log_num_blocks) { return log_num_blocks < 16 ? log_num_blocks : 16; }
inline int num_slices(int log_num_blocks) {
  int r = reachable_log_blocks(log_num_blocks);
  return r <= kSliceLog ? 1 : 1 << (r - kSliceLog);
}

// One key's test for member q of a group of G: true iff the member owns the
// key's block and the block holds every bit of the key's mask.
template <bool HAS_L2>
__device__ __forceinline__ bool find_test(uint32_t h, uint32_t bmask, uint32_t G, uint32_t q,
                                          const uint64_t* s_blk, const uint64_t* s_masks,
                                          const uint64_t* __restrict__ blocks) {
  const uint32_t b = (h >> (kMaskLogNum + 6)) & bmask;
  const uint32_t sl = b >> kSliceLog;
  bool mine = sl == q;
  uint64_t blk = s_blk[b & ((1u << kSliceLog) - 1)];
  if (HAS_L2) {
    const bool l2 = sl >= G && (b & (G - 1)) == q;
    const uint64_t g = blocks[l2 ? b : 0];
    blk = l2 ? g : blk;
    mine |= l2;
  }
  const uint64_t m = rotl64(s_masks[h & (kMaskNum - 1)], (h >> kMaskLogNum) & 63);
  return mine & ((~blk & m) == 0);
}

// PARTIAL: member q writes partial words to out_words[q * nwords + w];
// otherwise (G == 1) the bitmap is written directly (FindImp's tail-byte rule)
// and per-workgroup counts go to partials.
template <int MODE, bool HAS_L2, bool PARTIAL>
__global__ __launch_bounds__(kSliceThreads) void blocked_find_kernel(
    const uint64_t* __restrict__ blocks, uint64_t nb_mask, int group_size, ColSet keys, int64_t n,
    int64_t num_groups, uint64_t* __restrict__ out_words, uint8_t* __restrict__ bitmap,
    unsigned long long* partials) {
  extern __shared__ uint64_t lds[];
  uint64_t* s_masks = lds;
  uint64_t* s_blk = lds + kMaskNum;
  const uint32_t G = (uint32_t)group_size;
  uint32_t q;
  int64_t group;
  if (PARTIAL) {
    const uint32_t j = blockIdx.x >> 3;
    q = j % G;
    group = (int64_t)(j / G) * 8 + (blockIdx.x & 7);
  } else {
    q = 0;
    group = blockIdx.x;
  }
  const uint64_t reach = nb_mask + 1 < (1ULL << 16) ? nb_mask + 1 : (1ULL << 16);
  const uint64_t lo = (uint64_t)q << kSliceLog;
  const uint64_t cnt = reach - lo < (1ULL << kSliceLog) ? reach - lo : (1ULL << kSliceLog);
  for (int i = threadIdx.x; i < kMaskNum; i += blockDim.x) s_masks[i] = c_masks.m[i];
  stage_words(s_blk, blocks + lo, cnt);
  __syncthreads();
  const uint32_t bmask = (uint32_t)(reach - 1);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  const int64_t nwords = (n + 63) >> 6;
  const int64_t wfull = n >> 6;   // words whose 64 rows all exist
  // Steps of kSliceUnroll words are dealt round-robin over every group's
  // waves, so all workgroups stream through neighbouring stretches of the key
  // column at any moment (better HBM page locality than one distant range per
  // workgroup: 6.2 TB/s against 5.1 TB/s for a plain read on this part).
  const int64_t step = num_groups * nwaves * kSliceUnroll;
  uint64_t* __restrict__ mine = PARTIAL ? out_words + (int64_t)q * nwords : nullptr;
  uint64_t count = 0;
  auto emit = [&](int64_t w, uint64_t word) {
    if (PARTIAL) {
      if (lane == 0) mine[w] = word;
    } else {
      write_bitmap_word(bitmap, w, word, n, lane);
      count += __popcll(word);
    }
  };
  int64_t wb = (group * nwaves + wave) * kSliceUnroll;
  if (MODE == 1 || MODE == 2) {
    // whole steps of full words: no clamping, one address for the step's loads
    using T = typename std::conditional<MODE == 1, uint64_t, uint32_t>::type;
    const T* __restrict__ kp = reinterpret_cast<const T*>(keys.values[0]) + keys.offset[0] + lane;
    for (; wb + kSliceUnroll <= wfull; wb += step) {
      const T* p = kp + wb * 64;
      T x[kSliceUnroll];
#pragma unroll
      for (int u = 0; u < kSliceUnroll; ++u) x[u] = p[u * 64];
      uint64_t word[kSliceUnroll];
#pragma unroll
      for (int u = 0; u < kSliceUnroll; ++u)
        word[u] = __ballot(find_test<HAS_L2>(arrow6_hash_word(x[u]), bmask, G, q, s_blk, s_masks, blocks));
#pragma unroll
      for (int u = 0; u < kSliceUnroll; ++u) emit(wb + u, word[u]);
    }
  }
  // the rest (and every step of the generic key path): rows past the end
  // re-read the last row, so all of a step's loads are issued together
  const int64_t last_row = n - 1;
  for (; wb < nwords; wb += step) {
    uint32_t h[kSliceUnroll];
    bool in[kSliceUnroll];
#pragma unroll
    for (int u = 0; u < kSliceUnroll; ++u) {
      int64_t r = (wb + u) * 64 + lane;
      in[u] = r < n;
      h[u] = key_hash<MODE>(keys, in[u] ? r : last_row);
    }
    uint64_t word[kSliceUnroll];
#pragma unroll
    for (int u = 0; u < kSliceUnroll; ++u)
      word[u] = __ballot(in[u] & find_test<HAS_L2>(h[u], bmask, G, q, s_blk, s_masks, blocks));
#pragma unroll
    for (int u = 0; u < kSliceUnroll; ++u)
      if (wb + u < nwords) emit(wb + u, word[u]);
  }
  if (!PARTIAL && partials) block_count_store(partials, lane == 0 ? count : 0);
}

// bitmap word w = OR of the G partial words; FindImp's tail-byte rule for the
// last word; per-workgroup selected counts.
__global__ __launch_bounds__(256) void find_combine_kernel(const uint64_t* __restrict__ part, int slices,
                                                           int64_t n, uint8_t* __restrict__ bitmap,
                                                           unsigned long long* partials) {
  const int64_t nwords = (n + 63) >> 6;
  unsigned long long count = 0;
  for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nwords;
       w += (int64_t)gridDim.x * blockDim.x) {
    uint64_t word = 0;
    for (int q = 0; q < slices; ++q) word |= part[(int64_t)q * nwords + w];
    count += __popcll(word);
    if (w * 64 + 64 <= n) {
      reinterpret_cast<uint64_t*>(bitmap)[w] = word;
    } else {
      const int tail_bytes = (int)((n - w * 64 + 7) / 8);
      for (int i = 0; i < tail_bytes; ++i) bitmap[w * 8 + i] = (uint8_t)(word >> (8 * i));
    }
  }
  if (partials) block_count_store(partials, count);
}

template <bool HAS_L2, bool PARTIAL>
static void launch_find(int mode, dim3 g, size_t lds, hipStream_t s, const uint64_t* blocks, uint64_t nbm,
                        int G, const ColSet& k, int64_t n, int64_t ngroups, uint64_t* part, uint8_t* bitmap,
                        unsigned long long* parts) {
  dim3 b(kSliceThreads);
  switch (mode) {
    case 1: hipLaunchKernelGGL((blocked_find_kernel<1, HAS_L2, PARTIAL>), g, b, lds, s, blocks, nbm, G, k, n, ngroups, part, bitmap, parts); break;
    case 2: hipLaunchKernelGGL((blocked_find_kernel<2, HAS_L2, PARTIAL>), g, b, lds, s, blocks, nbm, G, k, n, ngroups, part, bitmap, parts); break;
    default: hipLaunchKernelGGL((blocked_find_kernel<0, HAS_L2, PARTIAL>), g, b, lds, s, blocks, nbm, G, k, n, ngroups, part, bitmap, parts); break;
  }
}

// Find of n rows into d_bitmap (and the selected count into d_num_selected
// when given).
static fpdb_status lds_find(Context* c, const uint64_t* d_blocks, int log_num_blocks, const ColSet& k, int64_t n,
                            uint8_t* d_bitmap, int64_t* d_num_selected) {
  const int S = num_slices(log_num_blocks);
  const int G = std::min(S, FPDB_FIND_GROUP);
  const bool has_l2 = G < S;
  const int64_t nwords = ceil_div(n, 64);
  const uint64_t nbm = (1ULL << log_num_blocks) - 1;
  const size_t lds = (size_t)(kMaskNum + (1 << std::min(kSliceLog, reachable_log_blocks(log_num_blocks)))) * 8;
  FPDB_CHECK(lds <= (size_t)c->max_lds_per_block, FPDB_ERR_INTERNAL, "bloom find: filter slice exceeds LDS");
  const int mode = key_mode(k);
  if (G == 1) {
    // two workgroups per CU when their LDS allows, contiguous word ranges
    const int per_cu = lds * 2 <= (size_t)c->max_lds_per_block ? 2 : 1;
    int64_t grid = std::max<int64_t>(1, (int64_t)std::max(c->num_cus, 1) * per_cu);
    grid = std::min<int64_t>(grid, ceil_div(nwords, kSliceUnroll * (kSliceThreads / 64)));
    FPDB_CHECK(grid <= kMaxPartials, FPDB_ERR_INTERNAL, "bloom find: too many workgroups");
    unsigned long long* parts = d_num_selected ? c->d_partials : nullptr;
    {
      ProfScope ps(c, "bloom_blocked_find");
      if (has_l2)
        launch_find<true, false>(mode, dim3((unsigned)grid), lds, c->stream, d_blocks, nbm, 1, k, n,
                                 grid, nullptr, d_bitmap, parts);
      else
        launch_find<false, false>(mode, dim3((unsigned)grid), lds, c->stream, d_blocks, nbm, 1, k, n,
                                  grid, nullptr, d_bitmap, parts);
    }
    FPDB_HIP(hipGetLastError());
    if (d_num_selected) FPDB_TRY(launch_sum_partials(c, (int)grid, d_num_selected));
    return FPDB_OK;
  }
  // one workgroup per CU (a slice fills most of the LDS); groups of G
  // workgroups, 8 groups (one per XCD) per 8 * G workgroups
  const int64_t per_round = 8LL * G;
  int64_t rounds = std::max<int64_t>(1, std::max(c->num_cus, 1) / per_round);
  rounds = std::max<int64_t>(1, std::min<int64_t>(rounds, ceil_div(nwords, 8LL * kSliceUnroll * (kSliceThreads / 64))));
  const int64_t grid = rounds * per_round;
  const int64_t ngroups = rounds * 8;
  uint64_t* part = nullptr;
  FPDB_TRY(ctx_alloc(c, (uint64_t)G * nwords * 8, reinterpret_cast<void**>(&part)));
  const int cgrid = grid_for(c, nwords, 256, 4);
  {
    ProfScope ps(c, "bloom_blocked_find");
    if (has_l2)
      launch_find<true, true>(mode, dim3((unsigned)grid), lds, c->stream, d_blocks, nbm, G, k, n, ngroups,
                              part, nullptr, nullptr);
    else
      launch_find<false, true>(mode, dim3((unsigned)grid), lds, c->stream, d_blocks, nbm, G, k, n, ngroups,
                               part, nullptr, nullptr);
    hipLaunchKernelGGL(find_combine_kernel, dim3(cgrid), dim3(256), 0, c->stream, part, G, n, d_bitmap,
                       d_num_selected ? c->d_partials : nullptr);
  }
  ctx_free(c, part);   // stream-ordered reuse: later work on the stream runs after these kernels
  FPDB_HIP(hipGetLastError());
  if (d_num_selected) FPDB_TRY(launch_sum_partials(c, cgrid, d_num_selected));
  return FPDB_OK;
}

// ---------------------------------------------------------------------------
// legacy filter
// ---------------------------------------------------------------------------

// Division of uint64 by an invariant divisor (Granlund & Montgomery 1994,
// "Division by invariant integers using multiplication", Fig. 4.1):
// q = SRL(t1 + SRL(n - t1, sh1), sh2), t1 = MULUH(mprime, n).
struct UDiv {
  uint64_t d;
  uint64_t mprime;
  int32_t sh1, sh2;
};

__host__ __device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

__host__ __device__ __forceinline__ uint64_t udiv_mod(uint64_t n, const UDiv& u) {
  uint64_t t1 = mulhi64(u.mprime, n);
  uint64_t q = (t1 + ((n - t1) >> u.sh1)) >> u.sh2;
  return n - q * u.d;
}

static UDiv make_udiv(uint64_t d) {
  UDiv u;
  u.d = d;
  int l = 0;
  while (l < 64 && (d > (1ULL << l))) ++l;    // l = ceil(log2 d)
  unsigned __int128 two_l = (unsigned __int128)1 << l;
  u.mprime = (uint64_t)((((unsigned __int128)1 << 64) * (two_l - d)) / d + 1);
  u.sh1 = l < 1 ? l : 1;
  u.sh2 = l - 1 > 0 ? l - 1 : 0;
  return u;
}

constexpr int kMaxLegacyK = 24;

struct LegacyParams {
  int64_t a[kMaxLegacyK];
  int64_t b[kMaxLegacyK];
  UDiv p[kMaxLegacyK];
  UDiv m[kMaxLegacyK];
  int32_t k;
};

// hashCombine over std::hash<intN> (identity) of each key column
// (fpdb-tuple/src/ArrayHasher.cpp:12-19, fpdb-util/src/Util.cpp:106-113).
// The legacy hasher reads the values buffer without looking at validity
// (ArrayHasher.cpp:47-53), so nulls hash their stored value.
__device__ __forceinline__ uint64_t legacy_key(const ColSet& k, int64_t row) {
  uint64_t seed = (uint64_t)k.n;
  for (int c = 0; c < k.n; ++c) {
    uint64_t h = (uint64_t)load_sext(k.values[c], k.width[c], row + k.offset[c]);
    seed ^= h + 0x9e3779b9ULL + (seed << 6) + (seed >> 2);
  }
  return seed;
}

// UniversalHashFunction::hash (UniversalHashFunction.cpp:51-57).
__device__ __forceinline__ uint64_t legacy_h(const LegacyParams& P, int j, uint64_t x) {
  uint64_t sum = (uint64_t)P.a[j] * x + (uint64_t)P.b[j];
  return udiv_mod(udiv_mod(sum, P.p[j]), P.m[j]);
}

__global__ __launch_bounds__(256) void legacy_build_kernel(ColSet keys, int64_t n,
                                                           const uint8_t* __restrict__ select,
                                                           LegacyParams P,
                                                           unsigned long long* __restrict__ words) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
    if (select && !bitmap_bit(select, r)) continue;
    uint64_t x = legacy_key(keys, r);
    for (int j = 0; j < P.k; ++j) {
      uint64_t h = legacy_h(P, j, x);
      __hip_atomic_fetch_or(&words[h >> 6], 1ULL << (h & 63), __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

__device__ __forceinline__ bool legacy_contains(const LegacyParams& P, const uint64_t* words,
                                                uint64_t x) {
  for (int j = 0; j < P.k; ++j) {
    uint64_t h = legacy_h(P, j, x);
    if (!((words[h >> 6] >> (h & 63)) & 1)) return false;
  }
  return true;
}

__global__ __launch_bounds__(256) void legacy_find_kernel(LegacyParams P, int64_t capacity,
                                                          const uint64_t* __restrict__ words,
                                                          ColSet keys, int64_t n,
                                                          uint8_t* __restrict__ bitmap,
                                                          unsigned long long* partials) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int64_t nchunks = (n + 63) >> 6;
  uint64_t count = 0;
  for (int64_t c = wave; c < nchunks; c += nwaves) {
    int64_t r = c * 64 + lane;
    bool sel = false;
    if (r < n && capacity != 0) sel = legacy_contains(P, words, legacy_key(keys, r));
    uint64_t word = __ballot(sel);
    write_bitmap_word(bitmap, c, word, n, lane);
    count += __popcll(word);
  }
  if (partials) block_count_store(partials, lane == 0 ? count : 0);
}

// column <op> constant -> selection bitmap (nulls never selected), optionally
// AND-ed (or OR-ed, for_or) with an existing bitmap.  Each wave reads and
// writes only its own 64-row words, so the existing bitmap may be the output.
__global__ __launch_bounds__(256) void compare_kernel(ColSet col, int32_t op, bool for_or, int64_t ci, double cf,
                                                      const uint8_t* and_bm, int64_t n, uint8_t* bitmap,
                                                      unsigned long long* partials) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int64_t nchunks = (n + 63) >> 6;
  const bool is_f = col.type[0] == FPDB_TYPE_FLOAT64 || col.type[0] == FPDB_TYPE_FLOAT32;
  const bool is_s = col.type[0] == FPDB_TYPE_INT8 || col.type[0] == FPDB_TYPE_INT16 ||
                    col.type[0] == FPDB_TYPE_INT32 || col.type[0] == FPDB_TYPE_INT64 ||
                    col.type[0] == FPDB_TYPE_DATE32 || col.type[0] == FPDB_TYPE_DATE64;
  uint64_t count = 0;
  for (int64_t c = wave; c < nchunks; c += nwaves) {
    int64_t r = c * 64 + lane;
    bool sel = false;
    if (r < n) {
      int64_t i = r + col.offset[0];
      int cmp;
      if (is_f) {
        double v = col.type[0] == FPDB_TYPE_FLOAT64 ? reinterpret_cast<const double*>(col.values[0])[i]
                                                   : (double)reinterpret_cast<const float*>(col.values[0])[i];
        cmp = v < cf ? -1 : (v > cf ? 1 : (v == cf ? 0 : 2));
      } else if (is_s) {
        int64_t v = load_sext(col.values[0], col.width[0], i);
        cmp = v < ci ? -1 : (v > ci ? 1 : 0);
      } else {
        uint64_t v = load_zext(col.values[0], col.width[0], i);
        cmp = v < (uint64_t)ci ? -1 : (v > (uint64_t)ci ? 1 : 0);
      }
      switch (op) {
        case FPDB_CMP_LT: sel = cmp == -1; break;
        case FPDB_CMP_LE: sel = cmp == -1 || cmp == 0; break;
        case FPDB_CMP_GT: sel = cmp == 1; break;
        case FPDB_CMP_GE: sel = cmp == 1 || cmp == 0; break;
        case FPDB_CMP_EQ: sel = cmp == 0; break;
        default: sel = cmp != 0; break;
      }
      sel = sel && valid_bit(col.validity[0], i);
      if (and_bm) sel = for_or ? (sel || bitmap_bit(and_bm, r)) : (sel && bitmap_bit(and_bm, r));
    }
    uint64_t word = __ballot(sel);
    write_bitmap_word(bitmap, c, word, n, lane);
    count += __popcll(word);
  }
  if (partials) block_count_store(partials, lane == 0 ? count : 0);
}

// blocked-filter predicate for the fused find + compaction

static fpdb_status check_outputs_match(const ColSet& in, const OutColSet& out) {
  FPDB_CHECK(in.n == out.n, FPDB_ERR_INVALID, "input and output column counts differ");
  for (int i = 0; i < in.n; ++i)
    FPDB_CHECK(in.width[i] == out.width[i], FPDB_ERR_INVALID,
               "output column " + std::to_string(i) + " width differs from input");
  return FPDB_OK;
}

}  // namespace fpdb

using namespace fpdb;

extern "C" {

fpdb_status fpdb_bloom_masks(uint8_t out_masks[136]) {
  FPDB_CHECK(out_masks, FPDB_ERR_INVALID, "NULL output");
  std::memcpy(out_masks, kBloomMaskBytes, kMaskTotalBytes);
  return FPDB_OK;
}

int fpdb_bloom_blocked_log_num_blocks(int64_t num_rows_to_insert) {
  int64_t desired = std::max<int64_t>(512, num_rows_to_insert * 8);
  int log_bits = 0;
  while ((int64_t(1) << log_bits) < desired) ++log_bits;
  return log_bits - 6;
}

fpdb_status fpdb_hash_keys(fpdb_context* ctx, const fpdb_column* keys, int num_keys,
                           int64_t num_rows, uint32_t* d_hashes) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  FPDB_CHECK(num_keys >= 1, FPDB_ERR_INVALID, "fpdb_hash_keys: no key columns");
  Context* c = as_ctx(ctx);
  ColSet k;
  FPDB_TRY(make_colset(keys, num_keys, num_rows, &k, "fpdb_hash_keys"));
  if (num_rows == 0) return FPDB_OK;
  hipLaunchKernelGGL(hash_keys_kernel, dim3(grid_for(c, num_rows, 256)), dim3(256), 0, c->stream, k,
                     num_rows, d_hashes);
  FPDB_HIP(hipGetLastError());
  return FPDB_OK;
}

static fpdb_status blocked_insert_impl(Context* c, const fpdb_column* keys, int num_keys,
                                       int64_t num_rows, const uint8_t* d_select,
                                       int log_num_blocks, uint64_t* d_blocks, bool clear) {
  FPDB_CHECK(num_keys >= 1, FPDB_ERR_INVALID, "bloom build: no key columns");
  FPDB_CHECK(log_num_blocks >= 0 && log_num_blocks <= 40, FPDB_ERR_INVALID, "bad log_num_blocks");
  FPDB_CHECK(d_blocks, FPDB_ERR_INVALID, "NULL blocks");
  ColSet k;
  FPDB_TRY(make_colset(keys, num_keys, num_rows, &k, "bloom build"));
  uint64_t nb = 1ULL << log_num_blocks;
  if (clear) FPDB_TRY(launch_fill(c, d_blocks, 0, nb * 8));
  if (num_rows == 0) return FPDB_OK;
  dim3 g(grid_for(c, num_rows, 256)), b(256);
  ProfScope ps(c, "bloom_blocked_insert");
  switch (key_mode(k)) {
    case 1: hipLaunchKernelGGL(blocked_insert_kernel<1>, g, b, 0, c->stream, k, num_rows, d_select, d_blocks, nb - 1); break;
    case 2: hipLaunchKernelGGL(blocked_insert_kernel<2>, g, b, 0, c->stream, k, num_rows, d_select, d_blocks, nb - 1); break;
    default: hipLaunchKernelGGL(blocked_insert_kernel<0>, g, b, 0, c->stream, k, num_rows, d_select, d_blocks, nb - 1); break;
  }
  FPDB_HIP(hipGetLastError());
  return FPDB_OK;
}

fpdb_status fpdb_bloom_blocked_build(fpdb_context* ctx, const fpdb_column* keys, int num_keys,
                                     int64_t num_rows, const uint8_t* d_select,
                                     int log_num_blocks, uint64_t* d_blocks) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  return blocked_insert_impl(as_ctx(ctx), keys, num_keys, num_rows, d_select, log_num_blocks, d_blocks, true);
}

fpdb_status fpdb_bloom_blocked_insert(fpdb_context* ctx, const fpdb_column* keys, int num_keys,
                                      int64_t num_rows, const uint8_t* d_select,
                                      int log_num_blocks, uint64_t* d_blocks) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  return blocked_insert_impl(as_ctx(ctx), keys, num_keys, num_rows, d_select, log_num_blocks, d_blocks, false);
}

fpdb_status fpdb_bloom_blocked_find(fpdb_context* ctx, const uint64_t* d_blocks, int log_num_blocks,
                                    const fpdb_column* keys, int num_keys, int64_t num_rows,
                                    uint8_t* d_bitmap, int64_t* d_num_selected) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  FPDB_CHECK(num_keys >= 1, FPDB_ERR_INVALID, "bloom find: no key columns");
  FPDB_CHECK(d_blocks && d_bitmap, FPDB_ERR_INVALID, "bloom find: NULL blocks or bitmap");
  FPDB_CHECK(log_num_blocks >= 0 && log_num_blocks <= 40, FPDB_ERR_INVALID, "bad log_num_blocks");
  Context* c = as_ctx(ctx);
  ColSet k;
  FPDB_TRY(make_colset(keys, num_keys, num_rows, &k, "bloom find"));
  if (num_rows == 0) return d_num_selected ? launch_fill(c, d_num_selected, 0, 8) : FPDB_OK;
  return lds_find(c, d_blocks, log_num_blocks, k, num_rows, d_bitmap, d_num_selected);
}

fpdb_status fpdb_bloom_legacy_params(int64_t capacity, double p, int64_t* k, int64_t* m) {
  FPDB_CHECK(k && m, FPDB_ERR_INVALID, "NULL output");
  FPDB_CHECK(p > 0.0 && p <= 1.0, FPDB_ERR_INVALID, "false positive rate must be in (0, 1]");
  FPDB_CHECK(capacity >= 0, FPDB_ERR_INVALID, "negative capacity");
  *k = (int64_t)std::ceil(std::log(1 / p) / std::log(2));
  *m = (int64_t)std::ceil(((double)capacity) * std::abs(std::log(p)) / std::pow(std::log(2), 2));
  return FPDB_OK;
}

static uint64_t mulmod_u64(uint64_t a, uint64_t b, uint64_t m) {
  return (uint64_t)((unsigned __int128)a * b % m);
}
static bool is_prime_u64(uint64_t n) {
  if (n < 2) return false;
  static const uint64_t bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  for (uint64_t q : bases) if (n % q == 0) return n == q;
  uint64_t d = n - 1;
  int s = 0;
  while (!(d & 1)) { d >>= 1; ++s; }
  for (uint64_t a : bases) {
    uint64_t x = 1, base = a % n, e = d;
    while (e) { if (e & 1) x = mulmod_u64(x, base, n); base = mulmod_u64(base, base, n); e >>= 1; }
    if (x == 1 || x == n - 1) continue;
    bool composite = true;
    for (int r = 1; r < s; ++r) {
      x = mulmod_u64(x, x, n);
      if (x == n - 1) { composite = false; break; }
    }
    if (composite) return false;
  }
  return true;
}

fpdb_status fpdb_bloom_legacy_make_hash_functions(uint32_t seed, int64_t num_bits, int k,
                                                  fpdb_universal_hash* out) {
  FPDB_CHECK(out && k >= 0 && k <= kMaxLegacyK, FPDB_ERR_INVALID, "bad arguments");
  FPDB_CHECK(num_bits > 0, FPDB_ERR_INVALID, "num_bits must be positive");
  std::mt19937 gen(seed);
  for (int j = 0; j < k; ++j) {
    int64_t i = std::uniform_int_distribution<int64_t>(num_bits, 2 * num_bits)(gen);
    uint64_t p = (uint64_t)i;
    while (!is_prime_u64(p)) ++p;
    int64_t a = std::uniform_int_distribution<int64_t>(1, (int64_t)p - 1)(gen);
    int64_t b = std::uniform_int_distribution<int64_t>(0, (int64_t)p - 1)(gen);
    out[j] = fpdb_universal_hash{a, b, num_bits, (int64_t)p};
  }
  return FPDB_OK;
}

static fpdb_status make_legacy_params(const fpdb_universal_hash* hf, int k, int64_t num_bits,
                                      LegacyParams* P) {
  FPDB_CHECK(k >= 0 && k <= kMaxLegacyK, FPDB_ERR_INVALID,
             "legacy bloom: k must be in [0, " + std::to_string(kMaxLegacyK) + "]");
  FPDB_CHECK(k == 0 || hf, FPDB_ERR_INVALID, "legacy bloom: NULL hash functions");
  std::memset(P, 0, sizeof(*P));
  P->k = k;
  for (int j = 0; j < k; ++j) {
    FPDB_CHECK(hf[j].p > 0 && hf[j].m > 0, FPDB_ERR_INVALID, "legacy bloom: p and m must be positive");
    FPDB_CHECK(num_bits < 0 || hf[j].m <= num_bits, FPDB_ERR_INVALID,
               "legacy bloom: hash function m exceeds the bit array");
    P->a[j] = hf[j].a;
    P->b[j] = hf[j].b;
    P->p[j] = make_udiv((uint64_t)hf[j].p);
    P->m[j] = make_udiv((uint64_t)hf[j].m);
  }
  return FPDB_OK;
}

fpdb_status fpdb_bloom_legacy_build(fpdb_context* ctx, const fpdb_column* keys, int num_keys,
                                    int64_t num_rows, const uint8_t* d_select,
                                    const fpdb_universal_hash* hash_functions, int k,
                                    int64_t num_bits, int64_t* d_words) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  FPDB_CHECK(num_keys >= 1, FPDB_ERR_INVALID, "legacy bloom: no key columns");
  FPDB_CHECK(num_bits >= 0 && d_words, FPDB_ERR_INVALID, "legacy bloom: bad bit array");
  Context* c = as_ctx(ctx);
  ColSet kc;
  FPDB_TRY(make_colset(keys, num_keys, num_rows, &kc, "legacy bloom build"));
  for (int i = 0; i < kc.n; ++i)
    FPDB_CHECK(type_is_signed_int(kc.type[i]) && (kc.width[i] == 4 || kc.width[i] == 8),
               FPDB_ERR_INVALID, "legacy bloom: key columns must be int32 or int64");
  LegacyParams P;
  FPDB_TRY(make_legacy_params(hash_functions, k, num_bits, &P));
  FPDB_TRY(launch_fill(c, d_words, 0, (uint64_t)(num_bits / 64 + 1) * 8));
  if (num_rows == 0 || k == 0) return FPDB_OK;
  ProfScope ps(c, "bloom_legacy_build");
  hipLaunchKernelGGL(legacy_build_kernel, dim3(grid_for(c, num_rows, 256)), dim3(256), 0, c->stream, kc,
                     num_rows, d_select, P, reinterpret_cast<unsigned long long*>(d_words));
  FPDB_HIP(hipGetLastError());
  return FPDB_OK;
}

fpdb_status fpdb_bloom_legacy_find(fpdb_context* ctx, const fpdb_universal_hash* hash_functions,
                                   int k, int64_t capacity, int64_t num_bits, const int64_t* d_words,
                                   const fpdb_column* keys, int num_keys, int64_t num_rows,
                                   uint8_t* d_bitmap, int64_t* d_num_selected) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  FPDB_CHECK(num_keys >= 1, FPDB_ERR_INVALID, "legacy bloom: no key columns");
  FPDB_CHECK(d_words && d_bitmap, FPDB_ERR_INVALID, "legacy bloom find: NULL buffer");
  Context* c = as_ctx(ctx);
  ColSet kc;
  FPDB_TRY(make_colset(keys, num_keys, num_rows, &kc, "legacy bloom find"));
  for (int i = 0; i < kc.n; ++i)
    FPDB_CHECK(type_is_signed_int(kc.type[i]) && (kc.width[i] == 4 || kc.width[i] == 8),
               FPDB_ERR_INVALID, "legacy bloom: key columns must be int32 or int64");
  FPDB_CHECK(num_bits >= 0, FPDB_ERR_INVALID, "legacy bloom find: negative num_bits");
  LegacyParams P;
  FPDB_TRY(make_legacy_params(hash_functions, k, num_bits, &P));
  if (num_rows == 0) return d_num_selected ? launch_fill(c, d_num_selected, 0, 8) : FPDB_OK;
  ProfScope ps(c, "bloom_legacy_find");
  const int grid = grid_for(c, num_rows, 256, 16);
  hipLaunchKernelGGL(legacy_find_kernel, dim3(grid), dim3(256), 0, c->stream, P, capacity,
                     reinterpret_cast<const uint64_t*>(d_words), kc, num_rows, d_bitmap,
                     d_num_selected ? c->d_partials : nullptr);
  FPDB_HIP(hipGetLastError());
  if (d_num_selected) FPDB_TRY(launch_sum_partials(c, grid, d_num_selected));
  return FPDB_OK;
}

fpdb_status fpdb_compare_bitmap(fpdb_context* ctx, const fpdb_column* column, int op, int64_t int_constant,
                                double float_constant, const uint8_t* d_and_bitmap, int64_t num_rows,
                                uint8_t* d_bitmap, int64_t* d_num_selected) {
  FPDB_CHECK(ctx && column && d_bitmap, FPDB_ERR_INVALID, "NULL argument");
  const bool for_or = (op & FPDB_CMP_OR) != 0;
  op &= ~FPDB_CMP_OR;
  FPDB_CHECK(op >= FPDB_CMP_LT && op <= FPDB_CMP_NE, FPDB_ERR_INVALID, "unknown comparison");
  FPDB_CHECK(!for_or || d_and_bitmap, FPDB_ERR_INVALID, "FPDB_CMP_OR needs a bitmap to combine with");
  Context* c = as_ctx(ctx);
  ColSet cs;
  FPDB_TRY(make_colset(column, 1, num_rows, &cs, "compare"));
  if (num_rows == 0) return d_num_selected ? launch_fill(c, d_num_selected, 0, 8) : FPDB_OK;
  ProfScope ps(c, "compare");
  const int grid = grid_for(c, num_rows, 256, 16);
  hipLaunchKernelGGL(compare_kernel, dim3(grid), dim3(256), 0, c->stream, cs, op, for_or, int_constant,
                     float_constant, d_and_bitmap, num_rows, d_bitmap, d_num_selected ? c->d_partials : nullptr);
  FPDB_HIP(hipGetLastError());
  if (d_num_selected) FPDB_TRY(launch_sum_partials(c, grid, d_num_selected));
  return FPDB_OK;
}

fpdb_status fpdb_filter_compact(fpdb_context* ctx, const fpdb_column* in_all, int num_all,
                                int64_t num_rows, const uint8_t* d_bitmap,
                                fpdb_column* out_all, int64_t* num_out) {
  FPDB_CHECK(ctx && num_out, FPDB_ERR_INVALID, "NULL argument");
  FPDB_CHECK(d_bitmap || num_rows == 0, FPDB_ERR_INVALID, "NULL bitmap");
  Context* c = as_ctx(ctx);
  std::vector<fpdb_column> fin, fout;
  std::vector<int> utf8;
  FPDB_TRY(split_utf8(in_all, out_all, num_all, &fin, &fout, &utf8, "filter compact"));
  ColSet ic;
  OutColSet oc;
  FPDB_TRY(make_colset(fin.data(), (int)fin.size(), num_rows, &ic, "filter compact"));
  FPDB_TRY(make_outcolset(fout.data(), (int)fout.size(), &oc, "filter compact"));
  FPDB_TRY(check_outputs_match(ic, oc));
  if (utf8.empty())
    return run_compact<PredBitmap, false>(c, PredBitmap{d_bitmap}, ic, oc, num_rows, nullptr, nullptr, num_out);
  // utf8 columns: the selected row numbers, then a take of each
  int64_t cap = oc.capacity;
  for (int i : utf8) {
    FPDB_CHECK(in_all[i].length >= num_rows, FPDB_ERR_INVALID, "filter compact: input column too short");
    cap = std::min<int64_t>(cap, out_all[i].length);
  }
  void* rows = nullptr;
  FPDB_TRY(ctx_alloc(c, (uint64_t)std::max<int64_t>(num_rows, 1) * 8, &rows));
  fpdb_status s = run_compact<PredBitmap, false>(c, PredBitmap{d_bitmap}, ic, oc, num_rows, nullptr,
                                                 static_cast<int64_t*>(rows), num_out, nullptr, "filter_compact", cap);
  for (size_t j = 0; s == FPDB_OK && j < utf8.size(); ++j)
    s = utf8_gather(c, in_all[utf8[j]], static_cast<int64_t*>(rows), *num_out, out_all[utf8[j]]);
  ctx_free(c, rows);
  return s;
}

fpdb_status fpdb_filter_compact_noalloc(fpdb_context* ctx, const fpdb_column* in, int num_columns,
                                        int64_t num_rows, const uint8_t* d_bitmap, fpdb_column* out,
                                        int64_t* d_num_out) {
  FPDB_CHECK(ctx && d_num_out, FPDB_ERR_INVALID, "NULL argument");
  FPDB_CHECK(d_bitmap || num_rows == 0, FPDB_ERR_INVALID, "NULL bitmap");
  FPDB_CHECK(num_rows >= 0 && num_rows <= kNoAllocMaxRows, FPDB_ERR_INVALID,
             "filter compact (no allocation): " + std::to_string(num_rows) +
                 " rows exceed the maximum of " + std::to_string(kNoAllocMaxRows) +
                 " (the library's static scratch); use fpdb_filter_compact");
  Context* c = as_ctx(ctx);
  FPDB_CHECK(c->noalloc_scratch, FPDB_ERR_INTERNAL, "no-allocation scratch address not set up");
  ColSet ic;
  OutColSet oc;
  FPDB_TRY(make_colset(in, num_columns, num_rows, &ic, "filter compact"));
  FPDB_TRY(make_outcolset(out, num_columns, &oc, "filter compact"));
  FPDB_TRY(check_outputs_match(ic, oc));
  FPDB_TRY(zero_out_validity(c, ic, oc, num_rows));
  if (num_rows == 0) return launch_fill(c, d_num_out, 0, 8);
  // scratch: tile counts (uint32) then tile bases (uint64, 8-byte aligned)
  const int64_t tiles = ceil_div(num_rows, kCompactTileRows);
  uint32_t* tcnt = reinterpret_cast<uint32_t*>(c->noalloc_scratch);
  uint64_t* tbase = reinterpret_cast<uint64_t*>(c->noalloc_scratch + (uint64_t)kNoAllocMaxTiles * 4);
  static_assert(kNoAllocMaxTiles * kCompactTileRows >= kNoAllocMaxRows, "scratch holds the tiles of the maximum");
  return compact_launch(c, PredBitmap{d_bitmap}, ic, oc, num_rows, nullptr, nullptr, tcnt, tbase, d_num_out,
                        oc.capacity, nullptr);
}

fpdb_status fpdb_bloom_blocked_filter(fpdb_context* ctx, const uint64_t* d_blocks, int log_num_blocks,
                                      const fpdb_column* in, int num_columns,
                                      const int* key_index, int num_keys, int64_t num_rows,
                                      uint8_t* d_bitmap, fpdb_column* out, int64_t* num_out) {
  FPDB_CHECK(ctx && num_out && d_blocks, FPDB_ERR_INVALID, "NULL argument");
  FPDB_CHECK(num_keys >= 1 && key_index, FPDB_ERR_INVALID, "bloom filter: no key columns");
  FPDB_CHECK(log_num_blocks >= 0 && log_num_blocks <= 40, FPDB_ERR_INVALID, "bad log_num_blocks");
  Context* c = as_ctx(ctx);
  ColSet ic;
  OutColSet oc;
  FPDB_TRY(make_colset(in, num_columns, num_rows, &ic, "bloom filter"));
  FPDB_TRY(make_outcolset(out, num_columns, &oc, "bloom filter"));
  FPDB_TRY(check_outputs_match(ic, oc));
  std::vector<fpdb_column> kcols;
  for (int j = 0; j < num_keys; ++j) {
    FPDB_CHECK(key_index[j] >= 0 && key_index[j] < num_columns, FPDB_ERR_INVALID, "bad key index");
    kcols.push_back(in[key_index[j]]);
  }
  ColSet kc;
  FPDB_TRY(make_colset(kcols.data(), num_keys, num_rows, &kc, "bloom filter keys"));
  if (num_rows == 0) return run_compact<PredBitmap, false>(c, PredBitmap{d_bitmap}, ic, oc, 0, nullptr, nullptr, num_out);
  // find with the filter in LDS, then an order-preserving compaction by the bitmap
  uint8_t* bm = d_bitmap;
  if (!bm) FPDB_TRY(ctx_alloc(c, (uint64_t)ceil_div(num_rows, 64) * 8, reinterpret_cast<void**>(&bm)));
  fpdb_status st = lds_find(c, d_blocks, log_num_blocks, kc, num_rows, bm, nullptr);
  if (st == FPDB_OK)
    st = run_compact<PredBitmap, false>(c, PredBitmap{bm}, ic, oc, num_rows, nullptr, nullptr, num_out, nullptr,
                                        "bloom_blocked_filter");
  if (!d_bitmap) ctx_free(c, bm);
  return st;
}

}  // extern "C"
