// Hash shuffle: the rows of a column set split into num_slots partitions by
// the Arrow 6 row hash of the key columns modulo num_slots, each partition
// keeping the input row order, written partition after partition
// (ShuffleKernel2::shuffle, fpdb-executor/src/physical/shuffle/ShuffleKernel2.cpp:28-126:
// row ids appended per partition at :66-69, concatenated at :85-89, Take at
// :94-97, sliced per partition at :113-119).
//
// A stable radix partition.  The rows are cut into tiles of kShufTileRows
// rows and the tiles into one contiguous run per workgroup (a few workgroups
// per compute unit, each looping over its run):
//   count    rows per partition of each run (LDS histogram), stored
//            partition-major: cnt[p * runs + g];
//   scan     one workgroup per partition scans its row of counts over the
//            runs, and one small kernel scans the partition totals into the
//            slot offsets;
//   scatter  each workgroup starts from its run's destination of every
//            partition and goes tile by tile: each row's rank among the tile's
//            rows of its partition (ballots over the partition-id bits give
//            the lanes of a wave with the same partition; running per-wave
//            counts in LDS order the chunks of a wave; a prefix over waves
//            orders the waves), then the rows are staged in LDS in partition
//            order and written out, so a wave's 64 stores go to one or two runs
//            of consecutive rows rather than 64 places; the destinations then
//            advance by the tile's counts.
// The key columns are read twice (count and scatter); every column is read
// once and written once; the per-partition bookkeeping in HBM is
// num_slots x runs words, not num_slots x tiles.
#include "compact.hpp"
#include "fpdb_internal.hpp"

namespace fpdb {

constexpr int kShufWaves = 4;
constexpr int kShufThreads = kShufWaves * 64;
#ifndef FPDB_SHUF_CHUNKS
#define FPDB_SHUF_CHUNKS 16
#endif
constexpr int kShufChunks = FPDB_SHUF_CHUNKS;                    // 64-row chunks per wave
constexpr int kShufTileRows = kShufWaves * kShufChunks * 64;     // 4096
constexpr int kShufRowsPerThread = kShufTileRows / kShufThreads; // 16
constexpr int kShufMaxSlots = 4096;
#ifndef FPDB_SHUF_RUNS_PER_CU
#define FPDB_SHUF_RUNS_PER_CU 0
#endif
constexpr int kShufRunsPerCu = FPDB_SHUF_RUNS_PER_CU;
constexpr int64_t kShufMaxCounts = int64_t(1) << 24;   // 16 Mi (slot, run) entries: 192 MiB of scratch
#ifndef FPDB_SHUF_MIN_WAVES
#define FPDB_SHUF_MIN_WAVES 3
#endif

// Lanes of the wave whose partition id equals this lane's (valid lanes only):
// one ballot per bit of the id.
__device__ __forceinline__ uint64_t match_partition(uint32_t p, bool valid, int bits) {
  uint64_t same = __ballot(valid);
  for (int b = 0; b < bits; ++b) {
    const bool bit = (p >> b) & 1;
    const uint64_t bal = __ballot(bit);
    same &= bit ? bal : ~bal;
  }
  return same;
}

// Partition ids of this lane's rows in the wave's chunks (a power-of-two slot
// count takes the low bits, which equals the modulo).
template <int MODE, bool POW2>
__device__ __forceinline__ void tile_partitions(const ColSet& keys, int64_t n, uint32_t num_slots,
                                                int64_t chunk0, int lane, uint32_t (&p)[kShufChunks]) {
#pragma unroll
  for (int c = 0; c < kShufChunks; ++c) {
    const int64_t row = (chunk0 + c) * 64 + lane;
    const uint32_t h = key_hash<MODE>(keys, row < n ? row : n - 1);
    p[c] = POW2 ? (h & (num_slots - 1)) : h % num_slots;
  }
}

// Tiles [t0, t1) of the workgroup's run (runs split the tiles evenly).
__device__ __forceinline__ void run_tiles(int64_t tiles, int64_t& t0, int64_t& t1) {
  t0 = tiles * (int64_t)blockIdx.x / gridDim.x;
  t1 = tiles * ((int64_t)blockIdx.x + 1) / gridDim.x;
}

// Count pass: rows per partition of run blockIdx.x into cnt[p * runs + g],
// and every row's partition id into pids (one or two bytes a row), so the
// scatter pass neither reads the keys again nor hashes them.
// One LDS atomic per row; with at most 4 slots, where a wave's atomics would
// mostly hit the same counter, the lanes of each partition are counted with
// ballots and one lane adds the count.
template <int MODE, bool POW2, class PID>
__global__ __launch_bounds__(kShufThreads) void shuffle_count_kernel(ColSet keys, int64_t n, uint32_t num_slots,
                                                                     int bits, int64_t tiles,
                                                                     uint32_t* __restrict__ cnt,
                                                                     PID* __restrict__ pids) {
  extern __shared__ uint64_t lds[];
  uint32_t* s_hist = reinterpret_cast<uint32_t*>(lds);
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  for (uint32_t q = threadIdx.x; q < num_slots; q += kShufThreads) s_hist[q] = 0;
  __syncthreads();
  int64_t t0, t1;
  run_tiles(tiles, t0, t1);
  for (int64_t tile = t0; tile < t1; ++tile) {
    const int64_t chunk0 = tile * (kShufWaves * kShufChunks) + wave * kShufChunks;
    uint32_t p[kShufChunks];
    tile_partitions<MODE, POW2>(keys, n, num_slots, chunk0, lane, p);
#pragma unroll
    for (int c = 0; c < kShufChunks; ++c) {
      const int64_t row = (chunk0 + c) * 64 + lane;
      if (row < n) pids[row] = (PID)p[c];
    }
    if (bits <= 2) {
#pragma unroll
      for (int c = 0; c < kShufChunks; ++c) {
        const bool valid = (chunk0 + c) * 64 + lane < n;
        const uint64_t same = match_partition(p[c], valid, bits);
        if (valid && rank_below(same) == 0) atomicAdd(&s_hist[p[c]], (uint32_t)__popcll(same));
      }
    } else {
#pragma unroll
      for (int c = 0; c < kShufChunks; ++c)
        if ((chunk0 + c) * 64 + lane < n) atomicAdd(&s_hist[p[c]], 1u);
    }
  }
  __syncthreads();
  for (uint32_t q = threadIdx.x; q < num_slots; q += kShufThreads)
    cnt[(int64_t)q * gridDim.x + blockIdx.x] = s_hist[q];
}

// Slot offsets: offsets[p] = rows of partitions below p, offsets[S] = all rows.
__global__ __launch_bounds__(1024) void shuffle_offsets_kernel(const int64_t* __restrict__ totals, int num_slots,
                                                               int64_t* __restrict__ offsets) {
  __shared__ int64_t s_w[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int per = (num_slots + 1023) / 1024;
  const int i0 = threadIdx.x * per;
  int64_t sum = 0;
  for (int i = 0; i < per; ++i)
    if (i0 + i < num_slots) sum += totals[i0 + i];
  int64_t incl = sum;
  for (int off = 1; off < 64; off <<= 1) {
    const int64_t u = __shfl_up(incl, off);
    if (lane >= off) incl += u;
  }
  if (lane == 63) s_w[wave] = incl;
  __syncthreads();
  int64_t before = incl - sum, all = 0;
  for (int w = 0; w < 16; ++w) {
    if (w < wave) before += s_w[w];
    all += s_w[w];
  }
  for (int i = 0; i < per; ++i) {
    if (i0 + i < num_slots) {
      offsets[i0 + i] = before;
      before += totals[i0 + i];
    }
  }
  if (threadIdx.x == 0) offsets[num_slots] = all;
}

// Exclusive scan of a[0, S) in LDS by the workgroup (values <= tile rows).
__device__ __forceinline__ void lds_exclusive_scan(uint32_t* a, uint32_t S, uint32_t* s_w) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t per = (S + kShufThreads - 1) / kShufThreads;
  const uint32_t i0 = threadIdx.x * per;
  uint32_t sum = 0;
  for (uint32_t i = 0; i < per; ++i)
    if (i0 + i < S) sum += a[i0 + i];
  uint32_t incl = sum;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t u = __shfl_up(incl, off);
    if (lane >= off) incl += u;
  }
  if (lane == 63) s_w[wave] = incl;
  __syncthreads();
  uint32_t before = incl - sum;
  for (int w = 0; w < wave; ++w) before += s_w[w];
  for (uint32_t i = 0; i < per; ++i) {
    if (i0 + i < S) {
      const uint32_t v = a[i0 + i];
      a[i0 + i] = before;
      before += v;
    }
  }
  __syncthreads();
}

template <int N>
__device__ __forceinline__ void store_w_n(uint8_t* base, int width, const int64_t (&idx)[N],
                                          const uint64_t (&v)[N], const bool (&on)[N]) {
  switch (width) {
    case 8: {
      uint64_t* q = reinterpret_cast<uint64_t*>(base);
#pragma unroll
      for (int i = 0; i < N; ++i)
        if (on[i]) q[idx[i]] = v[i];
      break;
    }
    case 4: {
      uint32_t* q = reinterpret_cast<uint32_t*>(base);
#pragma unroll
      for (int i = 0; i < N; ++i)
        if (on[i]) q[idx[i]] = (uint32_t)v[i];
      break;
    }
    case 2: {
      uint16_t* q = reinterpret_cast<uint16_t*>(base);
#pragma unroll
      for (int i = 0; i < N; ++i)
        if (on[i]) q[idx[i]] = (uint16_t)v[i];
      break;
    }
    default: {
#pragma unroll
      for (int i = 0; i < N; ++i)
        if (on[i]) base[idx[i]] = (uint8_t)v[i];
      break;
    }
  }
}

// LDS layout of the scatter kernel for S slots.
struct ShufLds {
  static constexpr size_t stage = 0;                                          // uint64[T]
  static constexpr size_t valid = stage + sizeof(uint64_t) * kShufTileRows;   // uint8[T]
  static constexpr size_t wsum = valid + kShufTileRows;                       // uint32[16]
  static constexpr size_t wcnt = wsum + 64;                                   // uint32[W * S]
  __host__ __device__ static size_t tstart(uint32_t S) { return wcnt + sizeof(uint32_t) * kShufWaves * S; }  // uint32[S]
  __host__ __device__ static size_t gbase(uint32_t S) { return (tstart(S) + sizeof(uint32_t) * S + 7) & ~size_t(7); }  // int64[S]
  __host__ __device__ static size_t bytes(uint32_t S) { return gbase(S) + sizeof(int64_t) * S; }
};

// Scatter pass over the tiles of run blockIdx.x.
template <class PID>
__global__ __launch_bounds__(kShufThreads, FPDB_SHUF_MIN_WAVES) void shuffle_scatter_kernel(
    const PID* __restrict__ pids, ColSet in, OutColSet out, int64_t n, uint32_t num_slots, int bits, int64_t tiles,
    const uint64_t* __restrict__ base, const int64_t* __restrict__ offsets) {
  extern __shared__ uint64_t lds[];
  uint8_t* L = reinterpret_cast<uint8_t*>(lds);
  uint64_t* s_stage = reinterpret_cast<uint64_t*>(L + ShufLds::stage);
  uint32_t* s_stage32 = reinterpret_cast<uint32_t*>(s_stage);
  uint8_t* s_valid = L + ShufLds::valid;
  uint32_t* s_wsum = reinterpret_cast<uint32_t*>(L + ShufLds::wsum);
  uint32_t* s_wcnt = reinterpret_cast<uint32_t*>(L + ShufLds::wcnt);
  uint32_t* s_tstart = reinterpret_cast<uint32_t*>(L + ShufLds::tstart(num_slots));
  int64_t* s_gbase = reinterpret_cast<int64_t*>(L + ShufLds::gbase(num_slots));

  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  uint32_t* my_cnt = s_wcnt + wave * num_slots;
  static_assert(kShufChunks * 64 <= 4096 && kShufMaxSlots <= (1 << 20), "packing");

  // destination of the run's first row of each partition
  for (uint32_t q = threadIdx.x; q < num_slots; q += kShufThreads)
    s_gbase[q] = offsets[q] + (int64_t)base[(int64_t)q * gridDim.x + blockIdx.x];
  int64_t t0, t1;
  run_tiles(tiles, t0, t1);
  for (int64_t tile = t0; tile < t1; ++tile) {
    const int64_t chunk0 = tile * (kShufWaves * kShufChunks) + wave * kShufChunks;
    const int64_t row0 = tile * kShufTileRows;
    const int tile_rows = n - row0 < kShufTileRows ? (int)(n - row0) : kShufTileRows;

    for (uint32_t i = threadIdx.x; i < kShufWaves * num_slots; i += kShufThreads) s_wcnt[i] = 0;
    uint32_t p[kShufChunks];
#pragma unroll
    for (int c = 0; c < kShufChunks; ++c) {
      const int64_t row = (chunk0 + c) * 64 + lane;
      p[c] = pids[row < n ? row : n - 1];
    }
    __syncthreads();

    // Rank of each row among the rows of its partition in this wave's
    // chunks, kept with the partition id as (p << 12) | rank.
#pragma unroll
    for (int c = 0; c < kShufChunks; ++c) {
      const bool valid = (chunk0 + c) * 64 + lane < n;
      const uint64_t same = match_partition(p[c], valid, bits);
      const uint32_t r = rank_below(same);
      const uint32_t prev = valid ? my_cnt[p[c]] : 0u;
      if (valid && r == 0) my_cnt[p[c]] = prev + (uint32_t)__popcll(same);
      p[c] = (p[c] << 12) | (prev + r);
    }
    __syncthreads();

    // Per partition: prefix over waves and the tile's total.
    for (uint32_t q = threadIdx.x; q < num_slots; q += kShufThreads) {
      uint32_t run = 0;
#pragma unroll
      for (int w = 0; w < kShufWaves; ++w) {
        const uint32_t t = s_wcnt[w * num_slots + q];
        s_wcnt[w * num_slots + q] = run;
        run += t;
      }
      s_tstart[q] = run;
    }
    __syncthreads();
    lds_exclusive_scan(s_tstart, num_slots, s_wsum);

    // Position of each row in the tile's partition order, and its
    // destination row (staged by position, then held by the thread that owns
    // the position; rows are numbered below 2^32).
    uint32_t pos[kShufChunks];
#pragma unroll
    for (int c = 0; c < kShufChunks; ++c) {
      const bool valid = (chunk0 + c) * 64 + lane < n;
      const uint32_t q = p[c] >> 12;
      const uint32_t within = my_cnt[q] + (p[c] & 4095u);
      pos[c] = valid ? s_tstart[q] + within : 0u;
      if (valid) s_stage32[pos[c]] = (uint32_t)(s_gbase[q] + within);
    }
    __syncthreads();
    uint32_t dst[kShufRowsPerThread];
#pragma unroll
    for (int k = 0; k < kShufRowsPerThread; ++k) {
      const int j = k * kShufThreads + threadIdx.x;
      dst[k] = s_stage32[j < tile_rows ? j : 0];
    }
    // the next tile's rows of each partition follow this tile's
    for (uint32_t q = threadIdx.x; q < num_slots; q += kShufThreads)
      s_gbase[q] += (q + 1 < num_slots ? s_tstart[q + 1] : (uint32_t)tile_rows) - s_tstart[q];
    __syncthreads();

  // Column by column: the rows go into LDS at their positions, then each
  // thread writes the positions it owns.  Half a wave's chunks at a time keeps
  // the registers for loads and stores in check.
  constexpr int H = kShufChunks / 2;
  const int64_t last = n - 1;
  for (int col = 0; col < in.n; ++col) {
    const int64_t ioff = in.offset[col];
    const bool nullable = in.validity[col] != nullptr;
#pragma unroll
    for (int c0 = 0; c0 < kShufChunks; c0 += H) {
      int64_t si[H];
      uint64_t v[H];
#pragma unroll
      for (int u = 0; u < H; ++u) {
        const int64_t r = (chunk0 + c0 + u) * 64 + lane;
        si[u] = (r < n ? r : last) + ioff;
      }
      load_zext_n<H>(in.values[col], in.width[col], si, v);
#pragma unroll
      for (int u = 0; u < H; ++u) {
        if ((chunk0 + c0 + u) * 64 + lane < n) {
          s_stage[pos[c0 + u]] = v[u];
          if (nullable) s_valid[pos[c0 + u]] = valid_bit(in.validity[col], si[u]);
        }
      }
    }
    __syncthreads();
    const int64_t ooff = out.offset[col];
#pragma unroll
    for (int k0 = 0; k0 < kShufRowsPerThread; k0 += H) {
      uint64_t w[H];
      int64_t di[H];
      bool o[H];
#pragma unroll
      for (int u = 0; u < H; ++u) {
        const int j = (k0 + u) * kShufThreads + threadIdx.x;
        o[u] = j < tile_rows;
        w[u] = s_stage[o[u] ? j : 0];
        di[u] = (int64_t)dst[k0 + u] + ooff;
      }
      store_w_n<H>(out.values[col], out.width[col], di, w, o);
      if (nullable) {
        uint32_t* ov = reinterpret_cast<uint32_t*>(out.validity[col]);
#pragma unroll
        for (int u = 0; u < H; ++u) {
          const int j = (k0 + u) * kShufThreads + threadIdx.x;
          if (o[u] && s_valid[j])
            __hip_atomic_fetch_or(ov + (di[u] >> 5), 1u << (di[u] & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    }
    __syncthreads();
  }
  }
}

static int bits_for(uint32_t num_slots) {
  int b = 0;
  while ((1u << b) < num_slots) ++b;
  return b;
}

template <class PID>
using ShufCountFn = void (*)(ColSet, int64_t, uint32_t, int, int64_t, uint32_t*, PID*);
template <class PID>
using ShufScatterFn = void (*)(const PID*, ColSet, OutColSet, int64_t, uint32_t, int, int64_t, const uint64_t*,
                               const int64_t*);

template <class PID>
static ShufCountFn<PID> shuffle_count_for(int mode, bool pow2) {
  switch (mode * 2 + (pow2 ? 1 : 0)) {
    case 2: return shuffle_count_kernel<1, false, PID>;
    case 3: return shuffle_count_kernel<1, true, PID>;
    case 4: return shuffle_count_kernel<2, false, PID>;
    case 5: return shuffle_count_kernel<2, true, PID>;
    case 1: return shuffle_count_kernel<0, true, PID>;
    default: return shuffle_count_kernel<0, false, PID>;
  }
}

// count -> scan -> offsets -> scatter for partition ids of type PID.
template <class PID>
static fpdb_status run_shuffle(Context* c, const ColSet& k, const ColSet& ic, const OutColSet& oc, int64_t n,
                               uint32_t S, size_t lds, int64_t* doff) {
  const int64_t tiles = ceil_div(n, kShufTileRows);
  const int bits = bits_for(S);
  const ShufCountFn<PID> count_fn = shuffle_count_for<PID>(key_mode(k), (S & (S - 1)) == 0);
  const ShufScatterFn<PID> scatter_fn = shuffle_scatter_kernel<PID>;
  // runs: one tile per workgroup (kShufRunsPerCu == 0, the measured best),
  // or kShufRunsPerCu x the scatter workgroups that stay resident at once.
  // The per-(slot, run) counts and bases are bounded by kShufMaxCounts
  // entries: beyond that the runs grow longer instead.
  int64_t runs = tiles;
  if (S * tiles > kShufMaxCounts) runs = std::max<int64_t>(1, kShufMaxCounts / S);
  if (kShufRunsPerCu > 0) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(scatter_fn),
                                                      kShufThreads, lds) != hipSuccess || per_cu < 1)
      per_cu = 1;
    runs = std::max<int64_t>(1, std::min<int64_t>(tiles, (int64_t)std::max(c->num_cus, 1) * per_cu * kShufRunsPerCu));
  }
  uint32_t* cnt = nullptr;
  uint64_t* tbase = nullptr;
  int64_t* totals = nullptr;
  PID* pids = nullptr;
  fpdb_status st = ctx_alloc(c, (uint64_t)S * runs * 4, reinterpret_cast<void**>(&cnt));
  if (st == FPDB_OK) st = ctx_alloc(c, (uint64_t)S * runs * 8, reinterpret_cast<void**>(&tbase));
  if (st == FPDB_OK) st = ctx_alloc(c, (uint64_t)S * 8, reinterpret_cast<void**>(&totals));
  if (st == FPDB_OK) st = ctx_alloc(c, (uint64_t)n * sizeof(PID), reinterpret_cast<void**>(&pids));
  if (st == FPDB_OK) {
    ProfScope ps(c, "shuffle");
    const dim3 grid((unsigned)runs), block(kShufThreads);
    {
      ProfScope ps1(c, "shuffle_count");
      hipLaunchKernelGGL(count_fn, grid, block, sizeof(uint32_t) * S, c->stream, k, n, S, bits, tiles, cnt, pids);
    }
    launch_tile_scan(c, S, cnt, runs, tbase, totals);
    hipLaunchKernelGGL(shuffle_offsets_kernel, dim3(1), dim3(1024), 0, c->stream, totals, (int)S, doff);
    {
      ProfScope ps2(c, "shuffle_scatter");
      hipLaunchKernelGGL(scatter_fn, grid, block, lds, c->stream, pids, ic, oc, n, S, bits, tiles, tbase, doff);
    }
    st = hipGetLastError() == hipSuccess ? FPDB_OK : fail(FPDB_ERR_HIP, "fpdb_shuffle: kernel launch failed");
  }
  // stream-ordered reuse: later work on the stream runs after these kernels
  if (cnt) ctx_free(c, cnt);
  if (tbase) ctx_free(c, tbase);
  if (totals) ctx_free(c, totals);
  if (pids) ctx_free(c, pids);
  return st;
}


}  // namespace fpdb

using namespace fpdb;

extern "C" fpdb_status fpdb_shuffle(fpdb_context* ctx, const fpdb_column* keys, int num_keys,
                                    const fpdb_column* in, int num_columns, int64_t num_rows, int num_slots,
                                    fpdb_column* out, int64_t* d_offsets, int64_t* offsets) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  FPDB_CHECK(num_keys >= 1, FPDB_ERR_INVALID, "fpdb_shuffle: no key columns");
  FPDB_CHECK(num_slots >= 1 && num_slots <= kShufMaxSlots, FPDB_ERR_INVALID,
             "fpdb_shuffle: num_slots must be in [1, " + std::to_string(kShufMaxSlots) + "]");
  FPDB_CHECK(num_rows >= 0 && num_rows < (int64_t(1) << 32), FPDB_ERR_INVALID,
             "fpdb_shuffle: row count must be in [0, 2^32)");
  FPDB_CHECK(num_columns >= 0 && (num_columns == 0 || (in && out)), FPDB_ERR_INVALID,
             "fpdb_shuffle: NULL columns");
  Context* c = as_ctx(ctx);
  ColSet k, ic;
  OutColSet oc;
  FPDB_TRY(make_colset(keys, num_keys, num_rows, &k, "fpdb_shuffle keys"));
  FPDB_TRY(make_colset(in, num_columns, num_rows, &ic, "fpdb_shuffle input"));
  FPDB_TRY(make_outcolset(out, num_columns, &oc, "fpdb_shuffle output"));
  for (int i = 0; i < num_columns; ++i)
    FPDB_CHECK(oc.width[i] == ic.width[i], FPDB_ERR_INVALID,
               "fpdb_shuffle: output column " + std::to_string(i) + " differs in width from its input");
  const uint32_t S = (uint32_t)num_slots;
  const size_t lds = ShufLds::bytes(S);
  FPDB_CHECK(lds <= std::min<size_t>(c->max_lds_per_block, 160 * 1024), FPDB_ERR_INVALID,
             "fpdb_shuffle: too many slots for LDS");
  FPDB_TRY(zero_out_validity(c, ic, oc, num_rows));

  int64_t* doff = d_offsets;
  if (!doff) FPDB_TRY(ctx_alloc(c, (uint64_t)(S + 1) * 8, reinterpret_cast<void**>(&doff)));
  fpdb_status st = FPDB_OK;
  if (num_rows == 0)
    st = launch_fill(c, doff, 0, (uint64_t)(S + 1) * 8);
  else if (S <= 256)
    st = run_shuffle<uint8_t>(c, k, ic, oc, num_rows, S, lds, doff);
  else
    st = run_shuffle<uint16_t>(c, k, ic, oc, num_rows, S, lds, doff);
  if (st == FPDB_OK && offsets) {
    hipError_t e = hipMemcpyAsync(offsets, doff, (uint64_t)(S + 1) * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) st = fail(FPDB_ERR_HIP, std::string("fpdb_shuffle: offsets read-back: ") + hipGetErrorString(e));
  }
  if (doff != d_offsets) ctx_free(c, doff);
  return st;
}
