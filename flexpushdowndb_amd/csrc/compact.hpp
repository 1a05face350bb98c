// Order-preserving stream compaction shared by the bloom-filter use path, the
// plain filter, and the build-side passes of outer/semi joins.
//
// A tile is 4 wavefronts x kCompactChunks chunks of 64 consecutive rows.  Each
// wave evaluates its predicate for 64 rows at a time and turns the results into
// one 64-bit word with a ballot (the LSB-first bitmap word of
// fpdb-tuple/src/arrow/exec/BloomFilter.cpp:138-170).  The predicates here are
// cheap to evaluate twice (a bitmap word, 64 flag bytes), so the compaction
// runs as a count pass (selected rows per tile), a scan of the tile counts and
// a copy pass that writes each selected row at its rank: the copy pass knows
// its offsets before it loads anything, so its loads are never held up by a
// chain of earlier tiles.
#pragma once

#include "fpdb_internal.hpp"
#include "bloom_masks.hpp"

#include <algorithm>
#include <type_traits>

namespace fpdb {

constexpr int kCompactWaves = 4;
#ifndef FPDB_COMPACT_CHUNKS
#define FPDB_COMPACT_CHUNKS 16   // 64-row chunks per wave (tile = 4 waves x this x 64 rows)
#endif
constexpr int kCompactChunks = FPDB_COMPACT_CHUNKS;
constexpr int64_t kCompactTileRows = kCompactWaves * kCompactChunks * 64;

__device__ __forceinline__ bool bitmap_bit(const uint8_t* bm, int64_t i) {
  return (bm[i >> 3] >> (i & 7)) & 1;
}

// Writes one 64-row bitmap word; a partial last word writes only
// ceil((n % 64) / 8) bytes (BloomFilter.cpp:167-169).
__device__ __forceinline__ void write_bitmap_word(uint8_t* bm, int64_t chunk, uint64_t word,
                                                  int64_t n, int lane) {
  int64_t row0 = chunk * 64;
  if (row0 + 64 <= n) {
    if (lane == 0) reinterpret_cast<uint64_t*>(bm)[chunk] = word;
  } else if (row0 < n) {
    int tail_bytes = (int)((n - row0 + 7) / 8);
    if (lane < tail_bytes) bm[chunk * 8 + lane] = (uint8_t)(word >> (8 * lane));
  }
}

// out.values[c][dst] = in.values[c][src]; validity bit OR-ed into a zeroed
// output bitmap.
__device__ __forceinline__ void copy_value(const ColSet& in, const OutColSet& out, int c,
                                           int64_t src, int64_t dst) {
  int64_t si = src + in.offset[c];
  uint64_t v = load_zext(in.values[c], in.width[c], si);
  int64_t di = dst + out.offset[c];
  store_w(out.values[c], out.width[c], di, v);
  if (out.validity[c] && valid_bit(in.validity[c], si))
    __hip_atomic_fetch_or(reinterpret_cast<uint32_t*>(out.validity[c]) + (di >> 5),
                          1u << (di & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Selection predicates of the compaction: the 64-bit selection word of rows
// [64 * chunk, 64 * chunk + 64), zero past row n (every lane gets the word).
// Bytes of a bitmap word that is partial (past row n) or unaligned, kept out
// of line so the unrolled callers stay small.
__device__ __noinline__ uint64_t bitmap_word_slow(const uint8_t* bm, int64_t chunk, int64_t n) {
  int64_t row0 = chunk * 64;
  if (row0 >= n) return 0;
  int nb = row0 + 64 <= n ? 8 : (int)((n - row0 + 7) / 8);
  uint64_t w = 0;
  for (int i = 0; i < nb; ++i) w |= (uint64_t)bm[chunk * 8 + i] << (8 * i);
  int rem = (int)(n - row0);
  return w & (rem >= 64 ? ~0ULL : ((1ULL << rem) - 1));
}

struct PredBitmap {
  const uint8_t* bm;   // LSB-first, ceil(n / 8) bytes
  __device__ __forceinline__ uint64_t word(int64_t chunk, int lane, int64_t n) const {
    if (chunk * 64 + 64 <= n && (reinterpret_cast<uintptr_t>(bm) & 7) == 0)
      return reinterpret_cast<const uint64_t*>(bm)[chunk];   // one broadcast load
    return bitmap_word_slow(bm, chunk, n);
  }
};

struct PredByteFlag {
  const uint8_t* flags;
  uint8_t want;
  __device__ __forceinline__ uint64_t word(int64_t chunk, int lane, int64_t n) const {
    int64_t r = chunk * 64 + lane;
    uint8_t f = flags[r < n ? r : n - 1];
    return __ballot(r < n && ((f != 0) == (want != 0)));
  }
};

static __constant__ ExpandedMasks c_masks = expand_masks();

// Exclusive scan of the tile counts (one workgroup of kScanThreads threads per
// segment of `tiles` counts: workgroup s scans cnt[s * tiles, (s + 1) * tiles)
// into the same range of base and stores its sum in total[s]).  Each thread
// owns kScanPerThread consecutive counts; global memory is read and written
// coalesced (consecutive threads, consecutive counts) and the counts change
// hands through LDS, since one workgroup's strided accesses would keep its
// CU's texture-address unit busy on one cache line per lane.  Inputs longer
// than kTileScanChunk counts are taken in chunks of that size, the next
// chunk's loads in flight while one is scanned.  (16 counts per thread, one chunk of
// 128 KiB for 16,384 tiles, measured slower on MI355X: 8.5 against 5.2 us for
// 5,859 tiles.)
constexpr int kScanThreads = 1024;
constexpr int kScanPerThread = 8;
constexpr int kTileScanChunk = kScanThreads * kScanPerThread;
static __global__ __launch_bounds__(kScanThreads) void tile_scan_kernel(const uint32_t* __restrict__ cnt,
                                                                        int64_t tiles, uint64_t* __restrict__ base,
                                                                        int64_t* total) {
  __shared__ __align__(16) uint64_t s_stage[kTileScanChunk];   // the chunk's counts (as uint32), then its bases
  __shared__ uint64_t s_w[kScanThreads / 64];
  uint32_t* s_cnt = reinterpret_cast<uint32_t*>(s_stage);
  cnt += (int64_t)blockIdx.x * tiles;
  base += (int64_t)blockIdx.x * tiles;
  total += blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  uint64_t carry = 0;
  uint32_t g[kScanPerThread];
  auto load_chunk = [&](int64_t c0) {
#pragma unroll
    for (int j = 0; j < kScanPerThread; ++j) {
      const int64_t i = c0 + j * kScanThreads + tid;
      g[j] = i < tiles ? cnt[i] : 0u;
    }
  };
  load_chunk(0);
  for (int64_t c0 = 0; c0 < tiles; c0 += kTileScanChunk) {
    const int len = tiles - c0 < kTileScanChunk ? (int)(tiles - c0) : kTileScanChunk;
#pragma unroll
    for (int j = 0; j < kScanPerThread; ++j) s_cnt[j * kScanThreads + tid] = g[j];
    __syncthreads();
    // the next chunk's counts are in flight while this one is scanned
    if (c0 + kTileScanChunk < tiles) load_chunk(c0 + kTileScanChunk);
    uint32_t v[kScanPerThread];
#pragma unroll
    for (int j = 0; j < kScanPerThread; j += 4) {
      const uint4 q = reinterpret_cast<const uint4*>(s_cnt)[(tid * kScanPerThread + j) >> 2];
      v[j] = q.x;
      v[j + 1] = q.y;
      v[j + 2] = q.z;
      v[j + 3] = q.w;
    }
    uint64_t mine = 0;
#pragma unroll
    for (int j = 0; j < kScanPerThread; ++j) mine += v[j];
    uint64_t incl = mine;
    for (int off = 1; off < 64; off <<= 1) {
      uint64_t u = __shfl_up(incl, off);
      if (lane >= off) incl += u;
    }
    if (lane == 63) s_w[wave] = incl;
    __syncthreads();   // also: every thread has read its counts from the stage
    uint64_t before = carry, chunk = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
      if (w < wave) before += s_w[w];
      chunk += s_w[w];
    }
    before += incl - mine;
#pragma unroll
    for (int j = 0; j < kScanPerThread; ++j) {
      s_stage[tid * kScanPerThread + j] = before;
      before += v[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kScanPerThread; ++j) {
      const int i = j * kScanThreads + tid;
      if (i < len) base[c0 + i] = s_stage[i];
    }
    carry += chunk;
    __syncthreads();   // the stage and s_w are reused by the next chunk
  }
  if (tid == 0) *total = (int64_t)carry;
}

// Scans `segments` segments of `tiles` counts each (see tile_scan_kernel).
inline void launch_tile_scan(Context* c, int segments, const uint32_t* cnt, int64_t tiles, uint64_t* base,
                             int64_t* total) {
  hipLaunchKernelGGL(tile_scan_kernel, dim3(segments), dim3(kScanThreads), 0, c->stream, cnt, tiles, base, total);
}

// Count pass: selected rows of each tile.
template <class Pred>
__global__ __launch_bounds__(256) void compact_count_kernel(Pred pred, int64_t n, uint32_t* __restrict__ tile_cnt) {
  __shared__ uint32_t s_w[kCompactWaves];
  const int64_t tile = blockIdx.x;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t wave_chunk0 = tile * (kCompactWaves * kCompactChunks) + wave * kCompactChunks;
  uint64_t words[kCompactChunks];
#pragma unroll
  for (int c = 0; c < kCompactChunks; ++c) words[c] = pred.word(wave_chunk0 + c, lane, n);
  uint32_t cnt = 0;
#pragma unroll
  for (int c = 0; c < kCompactChunks; ++c) cnt += __popcll(words[c]);
  if (lane == 0) s_w[wave] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < kCompactWaves; ++w) t += s_w[w];
    tile_cnt[tile] = t;
  }
}

// Count pass for a packed bitmap: a wave counts whole tiles, lane l holding
// word l of each 64 words of the tile, so a tile costs one coalesced load per
// 64 words; waves loop over tiles, kBitmapCountUnroll tiles in flight.  (One
// small workgroup per tile, as in compact_count_kernel, is bound by workgroup
// dispatch when each reads only a few hundred bytes.)
constexpr int kCompactTileWords = kCompactWaves * kCompactChunks;
static_assert(kCompactTileWords % 64 == 0, "a tile is whole wave-loads of bitmap words");
constexpr int kBitmapCountUnroll = 2;
static __global__ __launch_bounds__(256) void compact_count_bitmap_kernel(const uint8_t* __restrict__ bm, int64_t n,
                                                                          int64_t tiles,
                                                                          uint32_t* __restrict__ tile_cnt) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const bool aligned = (reinterpret_cast<uintptr_t>(bm) & 7) == 0;
  const int64_t full_words = n >> 6;   // words whose 64 rows all exist
  for (int64_t t0 = wave; t0 < tiles; t0 += nwaves * kBitmapCountUnroll) {
    uint32_t cnt[kBitmapCountUnroll];
#pragma unroll
    for (int u = 0; u < kBitmapCountUnroll; ++u) {
      const int64_t t = t0 + u * nwaves;
      cnt[u] = 0;
#pragma unroll
      for (int j = 0; j < kCompactTileWords; j += 64) {
        const int64_t w = t * kCompactTileWords + j + lane;
        // past the last row (t >= tiles included) the slow path reads nothing and gives 0
        const uint64_t word = aligned && w < full_words ? reinterpret_cast<const uint64_t*>(bm)[w]
                                                        : bitmap_word_slow(bm, w, n);
        cnt[u] += (uint32_t)__popcll(word);
      }
    }
#pragma unroll
    for (int u = 0; u < kBitmapCountUnroll; ++u) {
      uint32_t v = cnt[u];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      const int64_t t = t0 + u * nwaves;
      if (lane == 0 && t < tiles) tile_cnt[t] = v;
    }
  }
}

// Rank of this lane among the set bits of a wave-uniform word below it.
__device__ __forceinline__ uint32_t rank_below(uint64_t word) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(word >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)word, 0u));
}

__device__ __forceinline__ uint64_t wave_uniform(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Copy pass: each selected row of the tile written at tile_base[tile] + its
// rank in the tile; index_out (optional) gets the source row index of every
// selected row, bitmap_out (optional) the selection bitmap.  The selection
// words are wave-uniform and live in scalar registers; a lane's output
// position is the scalar running count plus an mbcnt.  Positions at or past
// `cap` (the rows the outputs hold) are not written; the caller compares the
// total with cap.
template <class Pred>
__global__ __launch_bounds__(256) void compact_copy_kernel(Pred pred, ColSet in, OutColSet out, int64_t n,
                                                           uint8_t* __restrict__ bitmap_out,
                                                           int64_t* __restrict__ index_out,
                                                           const uint64_t* __restrict__ tile_base, int64_t cap) {
  __shared__ uint32_t s_wave_cnt[kCompactWaves];
  const int64_t tile = blockIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int64_t wave_chunk0 = tile * (kCompactWaves * kCompactChunks) + wave * kCompactChunks;
  uint64_t words[kCompactChunks];
  uint32_t cnt = 0;
#pragma unroll
  for (int c = 0; c < kCompactChunks; ++c) words[c] = wave_uniform(pred.word(wave_chunk0 + c, lane, n));
#pragma unroll
  for (int c = 0; c < kCompactChunks; ++c) cnt += __popcll(words[c]);
  if (bitmap_out) {
#pragma unroll
    for (int c = 0; c < kCompactChunks; ++c)
      write_bitmap_word(bitmap_out, wave_chunk0 + c, words[c], n, lane);
  }
  if (lane == 0) s_wave_cnt[wave] = cnt;
  __syncthreads();
  uint64_t pos0 = tile_base[tile];
  for (int w = 0; w < wave; ++w) pos0 += s_wave_cnt[w];
  pos0 = wave_uniform(pos0);
  // Column by column, G chunks at a time: the G source values are loaded
  // together (rows past the end read the last row) before the selected lanes
  // store theirs, so no load waits in its own branch.
  constexpr int G = 8;
  static_assert(kCompactChunks % G == 0, "chunk groups");
  const int64_t last = n - 1;
  for (int col = 0; col < in.n; ++col) {
    uint64_t pos = pos0;
    const int64_t ioff = in.offset[col];
    const int64_t ooff = out.offset[col];
#pragma unroll
    for (int c0 = 0; c0 < kCompactChunks; c0 += G) {
      int64_t si[G];
      uint64_t v[G];
#pragma unroll
      for (int u = 0; u < G; ++u) {
        int64_t src = (wave_chunk0 + c0 + u) * 64 + lane;
        si[u] = (src < n ? src : last) + ioff;
      }
      load_zext_n<G>(in.values[col], in.width[col], si, v);
#pragma unroll
      for (int u = 0; u < G; ++u) {
        const uint64_t wc = words[c0 + u];
        if (((wc >> lane) & 1) && (int64_t)(pos + rank_below(wc)) < cap) {
          int64_t di = (int64_t)(pos + rank_below(wc)) + ooff;
          store_w(out.values[col], out.width[col], di, v[u]);
          if (out.validity[col] && valid_bit(in.validity[col], si[u]))
            __hip_atomic_fetch_or(reinterpret_cast<uint32_t*>(out.validity[col]) + (di >> 5), 1u << (di & 31),
                                  __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        pos += __popcll(wc);
      }
    }
  }
  if (index_out) {
    uint64_t pos = pos0;
#pragma unroll
    for (int c = 0; c < kCompactChunks; ++c) {
      const uint64_t wc = words[c];
      if (((wc >> lane) & 1) && (int64_t)(pos + rank_below(wc)) < cap)
        index_out[pos + rank_below(wc)] = (wave_chunk0 + c) * 64 + lane;
      pos += __popcll(wc);
    }
  }
}

inline int grid_for(Context* c, int64_t work_items, int per_block, int blocks_per_cu = 8) {
  int64_t need = ceil_div(std::max<int64_t>(work_items, 1), per_block);
  int64_t cap = (int64_t)std::max(c->num_cus, 1) * blocks_per_cu;
  return (int)std::max<int64_t>(1, std::min(need, cap));
}

// Zeroes output validity bitmaps (ceil(n/32) words each) that the kernel ORs into.
inline fpdb_status zero_out_validity(Context* c, const ColSet& in, const OutColSet& out, int64_t n) {
  for (int i = 0; i < out.n; ++i) {
    if (out.validity[i]) {
      FPDB_CHECK((out.offset[i] & 31) == 0, FPDB_ERR_INVALID,
                 "output validity offset must be a multiple of 32");
      FPDB_TRY(launch_fill(c, out.validity[i] + out.offset[i] / 8, 0, (uint64_t)((n + 31) / 32) * 4));
    }
    if (i < in.n && !out.validity[i] && in.validity[i])
      return fail(FPDB_ERR_INVALID, "input column " + std::to_string(i) +
                                        " has a validity bitmap but its output has none");
  }
  return FPDB_OK;
}

// The three launches of a compaction of n > 0 rows with caller-provided
// scratch: tile counts (tiles uint32) and tile bases (tiles uint64).  The
// selected count goes to the device int64 at `total`.  prof_name NULL: no
// timing scope (it would create events).
template <class Pred>
fpdb_status compact_launch(Context* c, const Pred& pred, const ColSet& in, const OutColSet& out, int64_t n,
                           uint8_t* bitmap_out, int64_t* index_out, uint32_t* tcnt, uint64_t* tbase,
                           int64_t* total, int64_t capacity, const char* prof_name) {
  const int64_t tiles = ceil_div(n, kCompactTileRows);
  auto launches = [&]() {
    if constexpr (std::is_same<Pred, PredBitmap>::value) {
      const int g = grid_for(c, ceil_div(tiles, kBitmapCountUnroll), 256 / 64, 8);
      hipLaunchKernelGGL(compact_count_bitmap_kernel, dim3(g), dim3(256), 0, c->stream, pred.bm, n, tiles, tcnt);
    } else {
      hipLaunchKernelGGL((compact_count_kernel<Pred>), dim3((unsigned)tiles), dim3(256), 0, c->stream, pred, n,
                         tcnt);
    }
    launch_tile_scan(c, 1, tcnt, tiles, tbase, total);
    hipLaunchKernelGGL((compact_copy_kernel<Pred>), dim3((unsigned)tiles), dim3(256), 0, c->stream, pred, in, out,
                       n, bitmap_out, index_out, tbase, capacity);
  };
  if (prof_name) {
    ProfScope ps(c, prof_name);
    launches();
  } else {
    launches();
  }
  FPDB_HIP(hipGetLastError());
  return FPDB_OK;
}

// Runs the compaction.  With d_num_out == nullptr the selected count is read
// back into *num_out (synchronizes); otherwise it is written to d_num_out on
// the device and the call returns without waiting.  Rows past `capacity`
// (the rows the output columns, and index_out, hold) are not written: the
// synchronous form then fails with FPDB_ERR_CAPACITY and *num_out holds the
// count the outputs need; the caller of the asynchronous form compares.
template <class Pred, bool NEED_MASKS>
fpdb_status run_compact(Context* c, const Pred& pred, const ColSet& in, const OutColSet& out,
                        int64_t n, uint8_t* bitmap_out, int64_t* index_out, int64_t* num_out,
                        int64_t* d_num_out = nullptr, const char* prof_name = "filter_compact",
                        int64_t capacity = INT64_MAX) {
  static_assert(!NEED_MASKS, "compaction predicates evaluate without the mask table");
  capacity = std::min(capacity, out.capacity);
  FPDB_TRY(zero_out_validity(c, in, out, n));
  if (n == 0) {
    if (d_num_out) return launch_fill(c, d_num_out, 0, 8);
    *num_out = 0;
    return FPDB_OK;
  }
  int64_t tiles = ceil_div(n, kCompactTileRows);
  FPDB_CHECK(tiles < (int64_t(1) << 31), FPDB_ERR_INVALID, "compaction: too many rows");
  uint32_t* tcnt = nullptr;
  uint64_t* tbase = nullptr;
  fpdb_status st = ctx_alloc(c, (uint64_t)tiles * 4, reinterpret_cast<void**>(&tcnt));
  if (st == FPDB_OK) st = ctx_alloc(c, (uint64_t)tiles * 8, reinterpret_cast<void**>(&tbase));
  if (st != FPDB_OK) {
    if (tcnt) ctx_free(c, tcnt);
    return st;
  }
  int64_t* total = d_num_out ? d_num_out : c->d_counter;
  st = compact_launch(c, pred, in, out, n, bitmap_out, index_out, tcnt, tbase, total, capacity, prof_name);
  ctx_free(c, tcnt);   // stream-ordered reuse: later work on the stream runs after these kernels
  ctx_free(c, tbase);
  FPDB_TRY(st);
  if (d_num_out) return FPDB_OK;
  FPDB_TRY(read_count(c, num_out));
  FPDB_CHECK(*num_out <= capacity, FPDB_ERR_CAPACITY,
             std::string(prof_name) + ": " + std::to_string(*num_out) + " rows selected, the output holds " +
                 std::to_string(capacity));
  return FPDB_OK;
}

}  // namespace fpdb
