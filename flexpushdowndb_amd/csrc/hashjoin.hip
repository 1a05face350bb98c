// Hash join build/probe on gfx950.
//
// Table: open addressing with linear probing in HBM, 16-byte slots
// {int64 key, uint32 build row, uint32 inline payload}, capacity a power of
// two >= 2x the build rows (load factor <= 0.5).  A probe therefore reads one
// 16-byte slot (one cache line) in the common case, and with an inline
// payload the build side needs no second gather.
//
// Build: one thread per build row claims a slot with a 32-bit CAS on the row
// field, then writes key and payload; keys are compared only by later probes.
//
// Probe: a match (or count) pass over tiles of 1024 probe rows records each
// row's first match and each tile's output count, a one-workgroup scan turns
// the counts into output offsets, and an emit pass writes probe columns and
// gathered build columns at their final positions, in probe-row order.
#include "fpdb_internal.hpp"
#include "compact.hpp"

#include <algorithm>
#include <vector>

namespace fpdb {

struct alignas(16) Slot {
  int64_t key;
  uint32_t row;
  uint32_t payload;
};
constexpr uint32_t kEmptyRow = 0xFFFFFFFFu;

// Canonical join key of up to two integer columns.
struct KeySpec {
  const uint8_t* v[2];
  const uint8_t* valid[2];
  int64_t off[2];
  int32_t w[2];
  int32_t sgn[2];
  int32_t n;
};

// Returns false for a null key.
__device__ __forceinline__ bool canon_key(const KeySpec& k, int64_t row, int64_t* out) {
  if (k.n == 1) {
    int64_t i = row + k.off[0];
    if (!valid_bit(k.valid[0], i)) return false;
    *out = k.sgn[0] ? load_sext(k.v[0], k.w[0], i) : (int64_t)load_zext(k.v[0], k.w[0], i);
    return true;
  }
  int64_t i0 = row + k.off[0], i1 = row + k.off[1];
  if (!valid_bit(k.valid[0], i0) || !valid_bit(k.valid[1], i1)) return false;
  uint32_t a = k.sgn[0] ? (uint32_t)load_sext(k.v[0], k.w[0], i0) : (uint32_t)load_zext(k.v[0], k.w[0], i0);
  uint32_t b = k.sgn[1] ? (uint32_t)load_sext(k.v[1], k.w[1], i1) : (uint32_t)load_zext(k.v[1], k.w[1], i1);
  *out = (int64_t)((uint64_t)a | ((uint64_t)b << 32));
  return true;
}

// KM: 1 = one non-null int64 column, 2 = one non-null int32 column, 0 = generic.
template <int KM>
__device__ __forceinline__ bool get_key(const KeySpec& k, int64_t row, int64_t* out) {
  if constexpr (KM == 1) {
    *out = reinterpret_cast<const int64_t*>(k.v[0])[row + k.off[0]];
    return true;
  } else if constexpr (KM == 2) {
    *out = reinterpret_cast<const int32_t*>(k.v[0])[row + k.off[0]];
    return true;
  } else {
    return canon_key(k, row, out);
  }
}

__device__ __forceinline__ uint64_t slot_of(int64_t key, uint64_t mask) {
  return fmix64((uint64_t)key) & mask;
}

__device__ __forceinline__ Slot load_slot(const Slot* slots, uint64_t s) {
  uint4 v = reinterpret_cast<const uint4*>(slots)[s];
  Slot r;
  r.key = (int64_t)((uint64_t)v.x | ((uint64_t)v.y << 32));
  r.row = v.z;
  r.payload = v.w;
  return r;
}

template <int KM>
__global__ __launch_bounds__(256) void hj_build_kernel(Slot* __restrict__ slots, uint64_t mask, KeySpec k,
                                                       int64_t n, const uint8_t* __restrict__ payload,
                                                       int64_t payload_off, int* err,
                                                       unsigned long long* partials) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  unsigned long long mine = 0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
    int64_t key;
    if (!get_key<KM>(k, r, &key)) continue;
    ++mine;
    uint64_t s = slot_of(key, mask);
    bool done = false;
    for (uint64_t probe = 0; probe <= mask; ++probe) {
      uint32_t old = atomicCAS(&slots[s].row, kEmptyRow, (uint32_t)r);
      if (old == kEmptyRow) {
        slots[s].key = key;
        slots[s].payload = payload ? reinterpret_cast<const uint32_t*>(payload)[r + payload_off] : 0u;
        done = true;
        break;
      }
      s = (s + 1) & mask;
    }
    if (!done) *err = 1;
  }
  block_count_store(partials, mine);   // inserted (non-null) rows
}

// Flags a table whose inserted keys are not all distinct.
__global__ void hj_unique_kernel(const Slot* __restrict__ slots, uint64_t mask, int* dup) {
  uint64_t cap = mask + 1;
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < cap;
       s += (uint64_t)gridDim.x * blockDim.x) {
    Slot me = load_slot(slots, s);
    if (me.row == kEmptyRow) continue;
    uint64_t h = slot_of(me.key, mask);
    while (h != s) {
      Slot o = load_slot(slots, h);
      if (o.key == me.key) { *dup = 1; break; }
      h = (h + 1) & mask;
    }
  }
}

// join kinds of the probe kernel
constexpr int kJtInner = 0;      // pairs
constexpr int kJtOuter = 1;      // pairs + unmatched probe rows
constexpr int kJtSemi = 2;       // probe rows with a match
constexpr int kJtMark = 3;       // no output, flag matched build rows

#ifndef FPDB_PROBE_ITEMS
#define FPDB_PROBE_ITEMS 4   // probe rows per thread (tile = 256 x this)
#endif
constexpr int kProbeItems = FPDB_PROBE_ITEMS;
constexpr int kProbeThreads = 256;
constexpr int kProbeWaves = kProbeThreads / 64;
constexpr int64_t kProbeTileRows = kProbeItems * kProbeThreads;

// Writes one joined row: probe columns copied from probe row r, build columns
// gathered from build row brow, or the inline payload, or (bit c of keymask:
// the column is the build key itself) the matching probe key.
__device__ __forceinline__ void emit_pair(const ColSet& pcols, const ColSet& bcols, const OutColSet& out,
                                          int64_t out_cap, int64_t r, int64_t dst, uint32_t brow,
                                          uint32_t payload, uint32_t keymask, int64_t key) {
  if (dst >= out_cap) return;
  for (int c = 0; c < pcols.n; ++c) copy_value(pcols, out, c, r, dst);
  for (int c = 0; c < bcols.n; ++c) {
    int oc = pcols.n + c;
    int64_t di = dst + out.offset[oc];
    bool valid;
    if (keymask & (1u << c)) {
      store_w(out.values[oc], out.width[oc], di, (uint64_t)key);
      valid = true;
    } else if (bcols.values[c] == nullptr) {
      store_w(out.values[oc], out.width[oc], di, payload);
      valid = true;
    } else {
      int64_t si = (int64_t)brow + bcols.offset[c];
      store_w(out.values[oc], out.width[oc], di, load_zext(bcols.values[c], bcols.width[c], si));
      valid = valid_bit(bcols.validity[c], si);
    }
    if (out.validity[oc] && valid)
      __hip_atomic_fetch_or(reinterpret_cast<uint32_t*>(out.validity[oc]) + (di >> 5), 1u << (di & 31),
                            __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// The probe runs as passes with known output offsets, so the random table
// reads never wait behind the output positions of earlier tiles:
//   match  tiles of 1024 rows (256 threads x 4): keys, table lookups (the home
//          slot and the next one loaded together), per row the match count and
//          first match (build row, inline payload), per tile the output count;
//   scan   exclusive scan of the tile counts (output offsets, total);
//   emit   the same tiles again: output positions from the tile offset and an
//          in-tile scan, then the output columns, written in probe-row order.
// Every phase issues its loads unconditionally (rows past the end read the
// last row) so none waits in its own branch.

// Keys of N rows row0 + i * stride (rows past the end read row n - 1 and are
// not live; a null key is not live).
template <int KM, int N, int STRIDE, bool NT = false>
__device__ __forceinline__ void probe_keys_n(const KeySpec& pk, int64_t n, int64_t row0, int64_t (&rc)[N],
                                             int64_t (&key)[N], bool (&live)[N]) {
#pragma unroll
  for (int i = 0; i < N; ++i) {
    int64_t r = row0 + (int64_t)i * STRIDE;
    rc[i] = r < n ? r : n - 1;
  }
  if (KM == 1 || KM == 2) {
    int64_t ki[N];
    uint64_t kv[N];
#pragma unroll
    for (int i = 0; i < N; ++i) ki[i] = rc[i] + pk.off[0];
    if (NT) load_zext_n_nt<N>(pk.v[0], KM == 1 ? 8 : 4, ki, kv);
    else load_zext_n<N>(pk.v[0], KM == 1 ? 8 : 4, ki, kv);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      key[i] = KM == 1 ? (int64_t)kv[i] : (int64_t)(int32_t)(uint32_t)kv[i];
      live[i] = row0 + (int64_t)i * STRIDE < n;
    }
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      key[i] = 0;
      live[i] = row0 + (int64_t)i * STRIDE < n && canon_key(pk, rc[i], &key[i]);
    }
  }
}

template <int KM>
__device__ __forceinline__ void probe_keys(const KeySpec& pk, int64_t n, int64_t row0, int64_t (&rc)[kProbeItems],
                                           int64_t (&key)[kProbeItems], bool (&live)[kProbeItems]) {
  probe_keys_n<KM, kProbeItems, kProbeThreads>(pk, n, row0, rc, key, live);
}

template <int JT>
__device__ __forceinline__ uint32_t probe_outputs(bool in_range, uint32_t cnt) {
  if (JT == kJtOuter) return in_range ? (cnt ? cnt : 1u) : 0u;
  if (JT == kJtSemi) return cnt ? 1u : 0u;
  if (JT == kJtMark) return 0u;
  return cnt;
}

template <int KM, int JT, bool TRACK>
__global__ __launch_bounds__(256) void hj_match_kernel(const Slot* __restrict__ slots, uint64_t mask, KeySpec pk,
                                                       int64_t n, uint8_t* __restrict__ matched, const int* dup_flag,
                                                       uint64_t* __restrict__ mrow, uint32_t* __restrict__ mcnt,
                                                       uint32_t* __restrict__ tile_cnt) {
  __shared__ uint32_t s_w[kProbeWaves];
  const int64_t tile = blockIdx.x;
  const int64_t row0 = tile * kProbeTileRows + threadIdx.x;
  // distinct build keys (checked when the table was built): a probe stops at
  // its first match
  const bool unique = *dup_flag == 0;
  int64_t key[kProbeItems], rc[kProbeItems];
  bool live[kProbeItems];
  probe_keys<KM>(pk, n, row0, rc, key, live);
  uint64_t home[kProbeItems];
  Slot first[kProbeItems], second[kProbeItems];
#pragma unroll
  for (int i = 0; i < kProbeItems; ++i) home[i] = slot_of(key[i], mask);
#pragma unroll
  for (int i = 0; i < kProbeItems; ++i) {
    first[i] = load_slot(slots, home[i]);
    second[i] = load_slot(slots, (home[i] + 1) & mask);
  }
  uint32_t cnt[kProbeItems], frow[kProbeItems], fpay[kProbeItems];
  uint64_t fslot[kProbeItems];
#pragma unroll
  for (int i = 0; i < kProbeItems; ++i) {
    cnt[i] = 0;
    frow[i] = 0;
    fpay[i] = 0;
    fslot[i] = 0;
    if (!live[i]) continue;
    Slot sl = first[i];
    uint64_t s = home[i];
    int step = 0;
    while (sl.row != kEmptyRow) {
      if (sl.key == key[i]) {
        if (cnt[i] == 0) { fslot[i] = s; frow[i] = sl.row; fpay[i] = sl.payload; }
        ++cnt[i];
        if (JT == kJtSemi || unique) break;
      }
      s = (s + 1) & mask;
      sl = (++step == 1) ? second[i] : load_slot(slots, s);
    }
  }
  if (TRACK) {
#pragma unroll
    for (int i = 0; i < kProbeItems; ++i) {
      if (!cnt[i]) continue;
      matched[frow[i]] = 1;   // flag every matched build row
      uint32_t seen = 1;
      uint64_t s = (fslot[i] + 1) & mask;
      while (seen < cnt[i]) {
        Slot sl = load_slot(slots, s);
        if (sl.row == kEmptyRow) break;
        if (sl.key == key[i]) { matched[sl.row] = 1; ++seen; }
        s = (s + 1) & mask;
      }
    }
  }
  uint32_t total = 0;
#pragma unroll
  for (int i = 0; i < kProbeItems; ++i) {
    int64_t r = row0 + (int64_t)i * kProbeThreads;
    if (r < n) {
      mcnt[r] = cnt[i];
      mrow[r] = ((uint64_t)frow[i] << 32) | fpay[i];
    }
    total += probe_outputs<JT>(r < n, cnt[i]);
  }
  for (int off = 32; off > 0; off >>= 1) total += __shfl_xor(total, off);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = total;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < kProbeWaves; ++w) t += s_w[w];
    tile_cnt[tile] = t;
  }
}

template <int KM, int JT>
__global__ __launch_bounds__(256) void hj_emit_kernel(const Slot* __restrict__ slots, uint64_t mask, KeySpec pk,
                                                      int64_t n, ColSet pcols, ColSet bcols, OutColSet out,
                                                      int64_t out_cap, uint32_t keymask,
                                                      const uint64_t* __restrict__ mrow,
                                                      const uint32_t* __restrict__ mcnt,
                                                      const uint64_t* __restrict__ tile_base) {
  __shared__ uint32_t s_cnt[kProbeItems][kProbeWaves];
  const int64_t tile = blockIdx.x;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t row0 = tile * kProbeTileRows + threadIdx.x;
  int64_t key[kProbeItems], rc[kProbeItems];
  bool live[kProbeItems];
  probe_keys<KM>(pk, n, row0, rc, key, live);
  uint32_t cnt[kProbeItems], frow[kProbeItems], fpay[kProbeItems], outc[kProbeItems];
#pragma unroll
  for (int i = 0; i < kProbeItems; ++i) {
    cnt[i] = mcnt[rc[i]];
    uint64_t m = mrow[rc[i]];
    frow[i] = (uint32_t)(m >> 32);
    fpay[i] = (uint32_t)m;
  }
#pragma unroll
  for (int i = 0; i < kProbeItems; ++i) {
    bool in_range = row0 + (int64_t)i * kProbeThreads < n;
    if (!in_range) cnt[i] = 0;
    outc[i] = probe_outputs<JT>(in_range, cnt[i]);
  }
  // positions within the tile, in row order (item-major, then thread)
  uint32_t incl[kProbeItems];
#pragma unroll
  for (int i = 0; i < kProbeItems; ++i) {
    incl[i] = wave_incl_scan(outc[i]);
    if (lane == 63) s_cnt[i][wave] = incl[i];
  }
  __syncthreads();
  uint64_t base = tile_base[tile];
  int64_t dst[kProbeItems];
  bool emit[kProbeItems];
#pragma unroll
  for (int i = 0; i < kProbeItems; ++i) {
    uint64_t before = base;
    for (int w = 0; w < wave; ++w) before += s_cnt[i][w];
    dst[i] = (int64_t)(before + incl[i] - outc[i]);
    for (int w = 0; w < kProbeWaves; ++w) base += s_cnt[i][w];
    emit[i] = outc[i] > 0 && dst[i] < out_cap;
  }
  // probe columns of the first output row: one column at a time, the values
  // of all items loaded together
  for (int c = 0; c < pcols.n; ++c) {
    int64_t si[kProbeItems];
    uint64_t v[kProbeItems];
#pragma unroll
    for (int i = 0; i < kProbeItems; ++i) si[i] = rc[i] + pcols.offset[c];
    load_zext_n<kProbeItems>(pcols.values[c], pcols.width[c], si, v);
#pragma unroll
    for (int i = 0; i < kProbeItems; ++i) {
      if (!emit[i]) continue;
      int64_t di = dst[i] + out.offset[c];
      store_w_nt(out.values[c], out.width[c], di, v[i]);
      if (out.validity[c] && valid_bit(pcols.validity[c], si[i]))
        __hip_atomic_fetch_or(reinterpret_cast<uint32_t*>(out.validity[c]) + (di >> 5), 1u << (di & 31),
                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (JT == kJtSemi) return;
  // build columns of the first output row: the first match, or null for an
  // unmatched outer row
  for (int c = 0; c < bcols.n; ++c) {
    const int oc = pcols.n + c;
    uint64_t v[kProbeItems];
    int64_t bi[kProbeItems];
#pragma unroll
    for (int i = 0; i < kProbeItems; ++i) bi[i] = (int64_t)(cnt[i] ? frow[i] : 0u) + bcols.offset[c];
    const bool from_key = (keymask >> c) & 1u;
    if (from_key) {
#pragma unroll
      for (int i = 0; i < kProbeItems; ++i) v[i] = (uint64_t)key[i];
    } else if (bcols.values[c] == nullptr) {
#pragma unroll
      for (int i = 0; i < kProbeItems; ++i) v[i] = fpay[i];
    } else {
      load_zext_n<kProbeItems>(bcols.values[c], bcols.width[c], bi, v);
    }
#pragma unroll
    for (int i = 0; i < kProbeItems; ++i) {
      if (!emit[i]) continue;
      int64_t di = dst[i] + out.offset[oc];
      bool m = cnt[i] != 0;
      store_w_nt(out.values[oc], out.width[oc], di, m ? v[i] : 0);
      bool valid = m && (from_key || bcols.values[c] == nullptr || valid_bit(bcols.validity[c], bi[i]));
      if (out.validity[oc] && valid)
        __hip_atomic_fetch_or(reinterpret_cast<uint32_t*>(out.validity[oc]) + (di >> 5), 1u << (di & 31),
                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  // later matches of a key (duplicate build keys): walk the chain again from
  // the home slot, skipping the first match
#pragma unroll
  for (int i = 0; i < kProbeItems; ++i) {
    if (cnt[i] < 2) continue;
    uint32_t seen = 0;
    uint64_t s = slot_of(key[i], mask);
    while (seen < cnt[i]) {
      Slot sl = load_slot(slots, s);
      if (sl.row == kEmptyRow) break;
      if (sl.key == key[i]) {
        if (seen > 0)
          emit_pair(pcols, bcols, out, out_cap, rc[i], dst[i] + seen, sl.row, sl.payload, keymask, key[i]);
        ++seen;
      }
      s = (s + 1) & mask;
    }
  }
}

// ---------------------------------------------------------------------------
// Dense (direct-index) tables
//
// A single integer key column whose non-null keys are distinct and span a
// range [kmin, kmax] of at most twice as many values as there are keys
// (surrogate keys, permutations) gets a collision-free table indexed by
// slot = key - kmin: build row (kEmptyRow where no key), inline payload, and a
// presence bitmap of one bit per slot.  No hashing, no chains.  The build runs
// without atomics in three launches: per-workgroup min/max partials together
// with the clear of every slot of the capacity; a plain scatter; a verify pass
// that flags a duplicate key (two rows scattered to one slot: one of them does
// not find itself there) together with the bitmap.  The scatter and verify
// workgroups each reduce the few minmax partials themselves, which costs less
// than a launch of its own.
struct DenseMeta {
  int64_t kmin, kmax, count, dup;
};

constexpr int kMinMaxThreads = 1024;
constexpr int kBuildUnroll = 4;   // rows a thread handles per trip in the dense build passes

// Per-workgroup {min, max, count} of the non-null build keys into partials
// (3 words per workgroup), and every slot of the capacity cleared (build row
// kEmptyRow, in the row array or in the row half of the {row, payload} pairs).
template <int KM>
__global__ __launch_bounds__(kMinMaxThreads) void hj_minmax_kernel(KeySpec k, int64_t n,
                                                                   int64_t* __restrict__ partials,
                                                                   uint32_t* __restrict__ rows,
                                                                   uint2* __restrict__ pairs, uint64_t cap,
                                                                   DenseMeta* meta) {
  const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t gthreads = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = gtid; i < cap; i += gthreads) {
    if (pairs)
      pairs[i] = make_uint2(kEmptyRow, 0u);
    else
      rows[i] = kEmptyRow;
  }
  if (gtid == 0) meta->dup = 0;
  int64_t lo = INT64_MAX, hi = INT64_MIN, cnt = 0;
  // kBuildUnroll rows a thread per trip, their keys loaded together
  for (int64_t r0 = (int64_t)gtid; r0 < n; r0 += (int64_t)gthreads * kBuildUnroll) {
    int64_t key[kBuildUnroll];
    bool ok[kBuildUnroll];
#pragma unroll
    for (int j = 0; j < kBuildUnroll; ++j) {
      const int64_t r = r0 + (int64_t)j * (int64_t)gthreads;
      key[j] = 0;
      ok[j] = get_key<KM>(k, r < n ? r : n - 1, &key[j]) && r < n;
    }
#pragma unroll
    for (int j = 0; j < kBuildUnroll; ++j) {
      if (!ok[j]) continue;
      lo = key[j] < lo ? key[j] : lo;
      hi = key[j] > hi ? key[j] : hi;
      ++cnt;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    int64_t l2 = __shfl_xor(lo, off), h2 = __shfl_xor(hi, off);
    lo = l2 < lo ? l2 : lo;
    hi = h2 > hi ? h2 : hi;
    cnt += __shfl_xor(cnt, off);
  }
  __shared__ int64_t s[3][kMinMaxThreads / 64];
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { s[0][wave] = lo; s[1][wave] = hi; s[2][wave] = cnt; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      lo = s[0][w] < lo ? s[0][w] : lo;
      hi = s[1][w] > hi ? s[1][w] : hi;
      cnt += s[2][w];
    }
    partials[3 * blockIdx.x] = lo;
    partials[3 * blockIdx.x + 1] = hi;
    partials[3 * blockIdx.x + 2] = cnt;
  }
}

struct DenseRange {
  int64_t kmin, kmax, count;
  uint64_t span;   // slots the keys span; 0 when there is no key or they do not fit cap slots
};

// The build keys' range from the minmax partials, reduced by the first wave of
// the calling workgroup and shared through LDS.  Every thread of the
// workgroup must call it (it synchronises the workgroup).
__device__ __forceinline__ DenseRange dense_range(const int64_t* __restrict__ partials, int parts, uint64_t cap) {
  __shared__ int64_t s_r[3];
  if (threadIdx.x < 64) {
    int64_t lo = INT64_MAX, hi = INT64_MIN, cnt = 0;
    for (int b = threadIdx.x; b < parts; b += 64) {
      const int64_t l = partials[3 * b], h = partials[3 * b + 1];
      lo = l < lo ? l : lo;
      hi = h > hi ? h : hi;
      cnt += partials[3 * b + 2];
    }
    for (int off = 32; off > 0; off >>= 1) {
      int64_t l2 = __shfl_xor(lo, off), h2 = __shfl_xor(hi, off);
      lo = l2 < lo ? l2 : lo;
      hi = h2 > hi ? h2 : hi;
      cnt += __shfl_xor(cnt, off);
    }
    if (threadIdx.x == 0) {
      s_r[0] = lo;
      s_r[1] = hi;
      s_r[2] = cnt;
    }
  }
  __syncthreads();
  DenseRange r;
  r.kmin = s_r[0];
  r.kmax = s_r[1];
  r.count = s_r[2];
  const uint64_t span = (uint64_t)r.kmax - (uint64_t)r.kmin + 1;
  r.span = (r.count > 0 && span != 0 && span <= cap) ? span : 0;
  return r;
}

// With an inline payload the scatter writes {row, payload} pairs into one
// 8-byte array (one random store per row instead of two) and
// hj_dense_finish_kernel then separates the two halves.
template <int KM>
__global__ __launch_bounds__(256) void hj_dense_scatter_kernel(uint32_t* __restrict__ rows, uint2* __restrict__ pairs,
                                                               uint64_t cap, const int64_t* __restrict__ partials,
                                                               int parts, KeySpec k, int64_t n,
                                                               const uint8_t* __restrict__ payload,
                                                               int64_t payload_off) {
  const DenseRange dr = dense_range(partials, parts, cap);
  if (dr.span == 0) return;
  const uint64_t kmin = (uint64_t)dr.kmin;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r0 < n; r0 += stride * kBuildUnroll) {
    int64_t key[kBuildUnroll];
    bool ok[kBuildUnroll];
    uint32_t pv[kBuildUnroll];
#pragma unroll
    for (int j = 0; j < kBuildUnroll; ++j) {
      const int64_t r = r0 + (int64_t)j * stride;
      const int64_t rc = r < n ? r : n - 1;
      key[j] = 0;
      ok[j] = get_key<KM>(k, rc, &key[j]) && r < n;
      pv[j] = pairs ? reinterpret_cast<const uint32_t*>(payload)[rc + payload_off] : 0u;
    }
#pragma unroll
    for (int j = 0; j < kBuildUnroll; ++j) {
      if (!ok[j]) continue;
      const int64_t r = r0 + (int64_t)j * stride;
      const uint64_t slot = (uint64_t)key[j] - kmin;
      if (pairs)
        pairs[slot] = make_uint2((uint32_t)r, pv[j]);
      else
        rows[slot] = (uint32_t)r;
    }
  }
}

// After the scatter, in one launch: the slots split into the row and payload
// arrays (when the scatter wrote pairs), the presence bitmap (bit s of word
// s / 64 set iff slot s holds a key), and per workgroup the number of filled
// slots, into filled[blockIdx.x].  Keys are distinct iff the filled slots
// number the keys (two rows of one key land in one slot), which
// hj_dense_decide_kernel checks; this needs no random read per build row.
// Workgroup 0 stores the range into meta for the host's one read-back.
template <int KM>
__global__ __launch_bounds__(256) void hj_dense_finish_kernel(const uint2* __restrict__ pairs, uint64_t cap,
                                                              const int64_t* __restrict__ partials, int parts,
                                                              DenseMeta* meta, uint32_t* __restrict__ rows,
                                                              uint32_t* __restrict__ pay, uint64_t* __restrict__ bits,
                                                              int64_t* __restrict__ filled) {
  const DenseRange dr = dense_range(partials, parts, cap);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    meta->kmin = dr.kmin;
    meta->kmax = dr.kmax;
    meta->count = dr.count;
  }
  const uint64_t span = dr.span;
  __shared__ unsigned long long s_fill;
  if (threadIdx.x == 0) s_fill = 0;
  __syncthreads();
  if (span == 0) {
    if (threadIdx.x == 0) filled[blockIdx.x] = 0;
    return;
  }
  const int lane = threadIdx.x & 63;
  const uint64_t gthreads = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // slots, a wave per 64 consecutive slots
  const uint64_t words = (span + 63) >> 6;
  const uint64_t wstride = gthreads >> 6;
  for (uint64_t w0 = gtid >> 6; w0 < words; w0 += wstride * kBuildUnroll) {
    uint2 e[kBuildUnroll];
#pragma unroll
    for (int j = 0; j < kBuildUnroll; ++j) {
      const uint64_t s = (w0 + j * wstride) * 64 + lane;
      const uint64_t sc = s < span ? s : 0;
      if (pairs) e[j] = pairs[sc];
      else e[j] = make_uint2(rows[sc], 0u);
    }
#pragma unroll
    for (int j = 0; j < kBuildUnroll; ++j) {
      const uint64_t w = w0 + j * wstride;
      const uint64_t s = w * 64 + lane;
      if (pairs && s < span) {
        rows[s] = e[j].x;
        pay[s] = e[j].y;
      }
      const uint64_t word = __ballot(s < span && e[j].x != kEmptyRow);
      if (lane == 0 && w < words) {
        bits[w] = word;
        atomicAdd(&s_fill, (unsigned long long)__popcll(word));
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) filled[blockIdx.x] = (int64_t)s_fill;
}

// One workgroup: the filled slots over the finish workgroups against the key
// count; meta->dup = 1 when they differ (a key occurs twice).
__global__ __launch_bounds__(1024) void hj_dense_decide_kernel(const int64_t* __restrict__ filled, int parts,
                                                               DenseMeta* meta) {
  __shared__ int64_t s_w[16];
  int64_t v = 0;
  for (int b = threadIdx.x; b < parts; b += blockDim.x) v += filled[b];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += s_w[w];
    meta->dup = t != meta->count ? 1 : 0;
  }
}

// Dense probe, pass 1: per probe row whether its key is present (the hit
// bitmap, 64 rows per word, LSB first) and per tile of kDenseTileRows rows its
// output count.  The presence bitmap is staged in LDS when it fits
// (BITS_IN_LDS), else read from L2.  Workgroups loop over tiles; one wave
// counts one tile at a time.  A wave loads its first tile's keys before the
// bitmap is staged and each next tile's keys before it tests the current one,
// so it waits for memory about once instead of once per tile.
constexpr int kDenseItems = 4;
constexpr int64_t kDenseTileRows = (int64_t)kDenseItems * kProbeThreads;
constexpr int kDensePrefetch = 4;   // probe columns loaded before the in-tile scan
constexpr int kDenseTileWords = (int)(kDenseTileRows / 64);
constexpr int kDenseCountThreads = 1024;

template <int KM, int JT, bool BITS_IN_LDS>
__global__ __launch_bounds__(kDenseCountThreads) void hj_dense_count_kernel(
    const uint64_t* __restrict__ bits, int64_t kmin, uint64_t span, KeySpec pk, int64_t n, int64_t tiles,
    uint64_t* __restrict__ hitmap, uint32_t* __restrict__ tile_cnt) {
  extern __shared__ uint64_t s_bits[];
  constexpr int U = kDenseTileWords;   // a tile's keys are all loaded together
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t)blockIdx.x * (kDenseCountThreads / 64) + (threadIdx.x >> 6);
  const int64_t nwaves = (int64_t)gridDim.x * (kDenseCountThreads / 64);
  int64_t key[U], rc[U];
  bool live[U];
  if (wave < tiles) probe_keys_n<KM, U, 64, true>(pk, n, wave * kDenseTileRows + lane, rc, key, live);
  const uint64_t nbw = (span + 63) >> 6;
  if (BITS_IN_LDS) {
    stage_words(s_bits, bits, nbw);
    __syncthreads();
  }
  const uint64_t* __restrict__ bw = BITS_IN_LDS ? s_bits : bits;
  for (int64_t t = wave; t < tiles; t += nwaves) {
    int64_t nkey[U], nrc[U];
    bool nlive[U];
    const int64_t tn = t + nwaves;
    if (tn < tiles) probe_keys_n<KM, U, 64, true>(pk, n, tn * kDenseTileRows + lane, nrc, nkey, nlive);
    // every lane reads its word (slot 0's for keys outside the span), so the
    // U reads are issued together
    uint64_t hw[U], pw[U];
    bool in[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t slot = (uint64_t)key[u] - (uint64_t)kmin;
      in[u] = live[u] && slot < span;
      const uint64_t s = in[u] ? slot : 0;
      pw[u] = bw[s >> 6] >> (s & 63);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) hw[u] = __ballot(in[u] & (bool)(pw[u] & 1));
    uint32_t cnt = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t w = t * kDenseTileWords + u;
      const int64_t r0 = w * 64;
      if (lane == 0 && r0 < n) hitmap[w] = hw[u];
      const int64_t rows_here = r0 >= n ? 0 : (n - r0 < 64 ? n - r0 : 64);
      if (JT == kJtOuter) cnt += (uint32_t)rows_here;
      else if (JT != kJtMark) cnt += (uint32_t)__popcll(hw[u]);
    }
    if (lane == 0) tile_cnt[t] = cnt;
    if (tn < tiles) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        key[u] = nkey[u];
        live[u] = nlive[u];
      }
    }
  }
}

// Dense probe, pass 2: a tile of kDenseTileRows probe rows (256 threads x 4,
// coalesced) writes its output rows at the scanned tile offset, in probe-row
// order.  The reads that do not depend on the scan -- keys, hit words, the
// slot's build row or payload, and the first kDensePrefetch probe columns --
// are all issued before the in-tile scan, so a wave waits for memory about
// twice instead of once per column.
// NEED_ROW: build columns gathered by build row, or build matches tracked.
template <int KM, int JT, bool TRACK, bool NEED_ROW>
__global__ __launch_bounds__(256) void hj_dense_emit_kernel(
    const uint32_t* __restrict__ drows, const uint32_t* __restrict__ dpay, int64_t kmin, uint64_t span, KeySpec pk,
    int64_t n, const uint64_t* __restrict__ hitmap, const uint64_t* __restrict__ tile_base, ColSet pcols,
    ColSet bcols, OutColSet out, int64_t out_cap, uint32_t keymask, uint8_t* __restrict__ matched) {
  __shared__ uint32_t s_cnt[kDenseItems][kProbeWaves];
  const int64_t tile = blockIdx.x;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t row0 = tile * kDenseTileRows + threadIdx.x;
  int64_t key[kDenseItems], rc[kDenseItems];
  bool live[kDenseItems];
  probe_keys_n<KM, kDenseItems, kProbeThreads, true>(pk, n, row0, rc, key, live);
  uint64_t hw[kDenseItems];
#pragma unroll
  for (int i = 0; i < kDenseItems; ++i) hw[i] = hitmap[rc[i] >> 6];
  // probe columns that are written whatever the scan says
  const bool emit_cols = !(JT == kJtMark || out.n == 0);
  // (a probe column that is the key column itself, non-null, takes the loaded
  // keys instead of a second load; the others load non-temporally)
  uint64_t pre[kDensePrefetch][kDenseItems];
#pragma unroll
  for (int c = 0; c < kDensePrefetch; ++c) {
    if (emit_cols && c < pcols.n) {
      if ((KM == 1 || KM == 2) && pcols.values[c] == pk.v[0] && pcols.offset[c] == pk.off[0] &&
          pcols.width[c] == (KM == 1 ? 8 : 4) && pcols.validity[c] == nullptr) {
#pragma unroll
        for (int i = 0; i < kDenseItems; ++i) pre[c][i] = KM == 1 ? (uint64_t)key[i] : (uint64_t)(uint32_t)key[i];
      } else {
        int64_t si[kDenseItems];
#pragma unroll
        for (int i = 0; i < kDenseItems; ++i) si[i] = rc[i] + pcols.offset[c];
        load_zext_n_nt<kDenseItems>(pcols.values[c], pcols.width[c], si, pre[c]);
      }
    }
  }
  uint32_t brow[kDenseItems], bpay[kDenseItems];
#pragma unroll
  for (int i = 0; i < kDenseItems; ++i) {
    const uint64_t slot = (uint64_t)key[i] - (uint64_t)kmin;
    const uint64_t s = slot < span ? slot : 0;
    brow[i] = NEED_ROW ? drows[s] : 0u;
    bpay[i] = dpay ? dpay[s] : 0u;
  }
  bool hit[kDenseItems];
  uint32_t outc[kDenseItems], incl[kDenseItems];
#pragma unroll
  for (int i = 0; i < kDenseItems; ++i) {
    hit[i] = live[i] && ((hw[i] >> (rc[i] & 63)) & 1);
    outc[i] = probe_outputs<JT>(row0 + (int64_t)i * kProbeThreads < n, hit[i] ? 1u : 0u);
    incl[i] = wave_incl_scan(outc[i]);
    if (lane == 63) s_cnt[i][wave] = incl[i];
  }
  if (TRACK) {
#pragma unroll
    for (int i = 0; i < kDenseItems; ++i)
      if (hit[i]) matched[brow[i]] = 1;
  }
  if (!emit_cols) return;
  __syncthreads();
  uint64_t base = tile_base[tile];
  int64_t dst[kDenseItems];
  bool emit[kDenseItems];
#pragma unroll
  for (int i = 0; i < kDenseItems; ++i) {
    uint64_t before = base;
    for (int w = 0; w < wave; ++w) before += s_cnt[i][w];
    dst[i] = (int64_t)(before + incl[i] - outc[i]);
    for (int w = 0; w < kProbeWaves; ++w) base += s_cnt[i][w];
    emit[i] = outc[i] > 0 && dst[i] < out_cap;
  }
  for (int c = 0; c < pcols.n; ++c) {
    int64_t si[kDenseItems];
    uint64_t v[kDenseItems];
#pragma unroll
    for (int i = 0; i < kDenseItems; ++i) si[i] = rc[i] + pcols.offset[c];
    if (c < kDensePrefetch) {
#pragma unroll
      for (int cc = 0; cc < kDensePrefetch; ++cc)
        if (cc == c) {
#pragma unroll
          for (int i = 0; i < kDenseItems; ++i) v[i] = pre[cc][i];
        }
    } else {
      load_zext_n<kDenseItems>(pcols.values[c], pcols.width[c], si, v);
    }
#pragma unroll
    for (int i = 0; i < kDenseItems; ++i) {
      if (!emit[i]) continue;
      int64_t di = dst[i] + out.offset[c];
      store_w(out.values[c], out.width[c], di, v[i]);
      if (out.validity[c] && valid_bit(pcols.validity[c], si[i]))
        __hip_atomic_fetch_or(reinterpret_cast<uint32_t*>(out.validity[c]) + (di >> 5), 1u << (di & 31),
                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (JT == kJtSemi) return;
  for (int c = 0; c < bcols.n; ++c) {
    const int oc = pcols.n + c;
    const bool from_key = (keymask >> c) & 1u;
    uint64_t v[kDenseItems];
    int64_t bi[kDenseItems];
#pragma unroll
    for (int i = 0; i < kDenseItems; ++i) bi[i] = (int64_t)(hit[i] ? brow[i] : 0u) + bcols.offset[c];
    if (from_key) {
#pragma unroll
      for (int i = 0; i < kDenseItems; ++i) v[i] = (uint64_t)key[i];
    } else if (bcols.values[c] == nullptr) {
#pragma unroll
      for (int i = 0; i < kDenseItems; ++i) v[i] = bpay[i];
    } else {
      load_zext_n<kDenseItems>(bcols.values[c], bcols.width[c], bi, v);
    }
#pragma unroll
    for (int i = 0; i < kDenseItems; ++i) {
      if (!emit[i]) continue;
      int64_t di = dst[i] + out.offset[oc];
      store_w(out.values[oc], out.width[oc], di, hit[i] ? v[i] : 0);
      bool valid = hit[i] && (from_key || bcols.values[c] == nullptr || valid_bit(bcols.validity[c], bi[i]));
      if (out.validity[oc] && valid)
        __hip_atomic_fetch_or(reinterpret_cast<uint32_t*>(out.validity[oc]) + (di >> 5), 1u << (di & 31),
                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

__global__ void gather_kernel(ColSet in, const int64_t* __restrict__ rows, int64_t n, OutColSet out) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    int64_t src = rows[i];
    for (int c = 0; c < in.n; ++c) {
      int64_t di = i + out.offset[c];
      if (src < 0) {
        store_w(out.values[c], out.width[c], di, 0);
        continue;
      }
      int64_t si = src + in.offset[c];
      store_w(out.values[c], out.width[c], di, load_zext(in.values[c], in.width[c], si));
      if (out.validity[c] && valid_bit(in.validity[c], si))
        __hip_atomic_fetch_or(reinterpret_cast<uint32_t*>(out.validity[c]) + (di >> 5),
                              1u << (di & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Copies n rows of `in` into `out` starting at output row `dst0`; validity bits
// are set or cleared individually, so any bit offset works.
__global__ void copy_rows_kernel(ColSet in, int64_t n, OutColSet out, int64_t dst0) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    for (int c = 0; c < in.n; ++c) {
      int64_t si = i + in.offset[c];
      int64_t di = dst0 + i + out.offset[c];
      store_w(out.values[c], out.width[c], di, load_zext(in.values[c], in.width[c], si));
      if (out.validity[c]) {
        uint32_t* w = reinterpret_cast<uint32_t*>(out.validity[c]) + (di >> 5);
        uint32_t bit = 1u << (di & 31);
        if (valid_bit(in.validity[c], si))
          __hip_atomic_fetch_or(w, bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else
          __hip_atomic_fetch_and(w, ~bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
}

}  // namespace fpdb

struct fpdb_hash_table {
  fpdb::Context* ctx = nullptr;
  fpdb::Slot* slots = nullptr;
  int64_t capacity = 0;
  int64_t build_rows = 0;
  int64_t inserted = 0;
  int num_keys = 0;
  int32_t key_types[2] = {0, 0};
  const void* key_values = nullptr;   // the build key column (compared, never read)
  int64_t key_offset = 0;
  bool has_payload = false;
  int32_t payload_type = 0;
  bool unique = true;
  bool info_ready = false;
  uint8_t* matched = nullptr;
  int64_t* meta = nullptr;    // device: failure flag, duplicate flag, inserted rows
  // dense (direct-index) form, slot = key - dense_kmin < dense_span: build
  // row per slot (kEmptyRow where no key), inline payload per slot, presence
  // bitmap; slots == nullptr then
  uint32_t* dense = nullptr;        // rows, then payloads (when has_payload)
  uint32_t* dense_pay = nullptr;
  uint64_t* dense_bits = nullptr;
  int64_t dense_kmin = 0;
  uint64_t dense_span = 0;
};

namespace fpdb {

static fpdb_status make_keyspec(const fpdb_column* keys, int num_keys, int64_t num_rows, KeySpec* k,
                                const char* what) {
  FPDB_CHECK(num_keys == 1 || num_keys == 2, FPDB_ERR_INVALID,
             std::string(what) + ": one or two key columns are supported");
  FPDB_CHECK(keys, FPDB_ERR_INVALID, std::string(what) + ": NULL keys");
  std::memset(k, 0, sizeof(*k));
  k->n = num_keys;
  for (int i = 0; i < num_keys; ++i) {
    int w = type_width(keys[i].type);
    FPDB_CHECK(w > 0 && !type_is_float(keys[i].type), FPDB_ERR_INVALID,
               std::string(what) + ": key columns must be integers");
    FPDB_CHECK(num_keys == 1 || w <= 4, FPDB_ERR_INVALID,
               std::string(what) + ": two-column keys need columns of at most 4 bytes");
    FPDB_CHECK(keys[i].length >= num_rows, FPDB_ERR_INVALID, std::string(what) + ": key column too short");
    FPDB_CHECK(num_rows == 0 || keys[i].values, FPDB_ERR_INVALID, std::string(what) + ": NULL key values");
    k->v[i] = static_cast<const uint8_t*>(keys[i].values);
    k->valid[i] = keys[i].validity;
    k->off[i] = keys[i].offset;
    k->w[i] = w;
    k->sgn[i] = type_is_signed_int(keys[i].type) ? 1 : 0;
  }
  return FPDB_OK;
}

static int key_mode_of(const KeySpec& k) {
  if (k.n == 1 && k.valid[0] == nullptr && k.sgn[0]) {
    if (k.w[0] == 8) return 1;
    if (k.w[0] == 4) return 2;
  }
  return 0;
}

static fpdb_status check_key_compat(const fpdb_hash_table* t, const fpdb_column* keys, int num_keys) {
  FPDB_CHECK(num_keys == t->num_keys, FPDB_ERR_INVALID, "probe key count differs from the build side");
  for (int i = 0; i < num_keys; ++i) {
    bool bs = type_is_signed_int(t->key_types[i]), ps = type_is_signed_int(keys[i].type);
    FPDB_CHECK(bs == ps, FPDB_ERR_INVALID, "probe and build key signedness differ");
  }
  return FPDB_OK;
}

// Dense probe: count pass (hit bitmap + tile counts), scan, emit pass.
static fpdb_status dense_probe(Context* c, fpdb_hash_table* t, int jt, bool track, const KeySpec& pk, int64_t n,
                               const ColSet& pc, const ColSet& bc, const OutColSet& oc, int64_t out_cap,
                               uint32_t keymask, int64_t* num_out, int64_t* d_num_out) {
  const int64_t tiles = ceil_div(n, kDenseTileRows);
  FPDB_CHECK(tiles < (int64_t(1) << 31), FPDB_ERR_INVALID, "hash join probe: too many rows");
  const int64_t hwords = tiles * kDenseTileWords;
  uint64_t* hitmap = nullptr;
  uint32_t* tcnt = nullptr;
  uint64_t* tbase = nullptr;
  fpdb_status st = ctx_alloc(c, (uint64_t)hwords * 8, reinterpret_cast<void**>(&hitmap));
  if (st == FPDB_OK) st = ctx_alloc(c, (uint64_t)tiles * 4, reinterpret_cast<void**>(&tcnt));
  if (st == FPDB_OK) st = ctx_alloc(c, (uint64_t)tiles * 8, reinterpret_cast<void**>(&tbase));
  auto release = [&]() {
    if (hitmap) ctx_free(c, hitmap);
    if (tcnt) ctx_free(c, tcnt);
    if (tbase) ctx_free(c, tbase);
  };
  if (st != FPDB_OK) {
    release();
    return st;
  }
  int64_t* total = d_num_out ? d_num_out : c->d_counter;
  const int km = key_mode_of(pk);
  const uint64_t span = t->dense_span;
  const size_t bits_bytes = ((span + 63) / 64) * 8;
  const bool in_lds = bits_bytes <= std::min<size_t>(c->max_lds_per_block, 160 * 1024);
  const int cgrid = (int)std::max<int64_t>(1, std::min<int64_t>(std::max(c->num_cus, 1),
                                                                 ceil_div(tiles, kDenseCountThreads / 64)));
  const bool need_row = track || [&]() {
    for (int i = 0; i < bc.n; ++i)
      if (!((keymask >> i) & 1u) && bc.values[i] != nullptr) return true;
    return false;
  }();
  const uint32_t* dpay = t->dense_pay;
  {
    ProfScope ps(c, "hj_probe");
    dim3 cg(cgrid), cb(kDenseCountThreads);
    const size_t lds = in_lds ? bits_bytes : 0;
#define FPDB_DCOUNT(KM, JT)                                                                                    \
  do {                                                                                                         \
    if (in_lds)                                                                                                \
      hipLaunchKernelGGL((hj_dense_count_kernel<KM, JT, true>), cg, cb, lds, c->stream, t->dense_bits,         \
                         t->dense_kmin, span, pk, n, tiles, hitmap, tcnt);                                     \
    else                                                                                                       \
      hipLaunchKernelGGL((hj_dense_count_kernel<KM, JT, false>), cg, cb, 0, c->stream, t->dense_bits,          \
                         t->dense_kmin, span, pk, n, tiles, hitmap, tcnt);                                     \
  } while (0)
#define FPDB_DCOUNT_JT(KM)                                 \
  switch (jt) {                                            \
    case kJtInner: FPDB_DCOUNT(KM, kJtInner); break;       \
    case kJtOuter: FPDB_DCOUNT(KM, kJtOuter); break;       \
    case kJtSemi: FPDB_DCOUNT(KM, kJtSemi); break;         \
    default: FPDB_DCOUNT(KM, kJtMark); break;              \
  }
    switch (km) {
      case 1: FPDB_DCOUNT_JT(1); break;
      case 2: FPDB_DCOUNT_JT(2); break;
      default: FPDB_DCOUNT_JT(0); break;
    }
#undef FPDB_DCOUNT_JT
#undef FPDB_DCOUNT
    launch_tile_scan(c, 1, tcnt, tiles, tbase, total);
    if ((jt != kJtMark && oc.n > 0 && out_cap > 0) || track) {
      dim3 g((unsigned)tiles), b(kProbeThreads);
#define FPDB_DEMIT(KM, JT, TR, NR)                                                                             \
  hipLaunchKernelGGL((hj_dense_emit_kernel<KM, JT, TR, NR>), g, b, 0, c->stream, t->dense, dpay, t->dense_kmin, \
                     span, pk, n, hitmap, tbase, pc, bc, oc, out_cap, keymask, t->matched)
#define FPDB_DEMIT_JT(KM)                                                         \
  switch (jt * 4 + (track ? 2 : 0) + (need_row ? 1 : 0)) {                       \
    case kJtInner * 4: FPDB_DEMIT(KM, kJtInner, false, false); break;             \
    case kJtInner * 4 + 1: FPDB_DEMIT(KM, kJtInner, false, true); break;          \
    case kJtInner * 4 + 3: FPDB_DEMIT(KM, kJtInner, true, true); break;           \
    case kJtOuter * 4: FPDB_DEMIT(KM, kJtOuter, false, false); break;             \
    case kJtOuter * 4 + 1: FPDB_DEMIT(KM, kJtOuter, false, true); break;          \
    case kJtOuter * 4 + 3: FPDB_DEMIT(KM, kJtOuter, true, true); break;           \
    case kJtSemi * 4: FPDB_DEMIT(KM, kJtSemi, false, false); break;               \
    case kJtSemi * 4 + 1: FPDB_DEMIT(KM, kJtSemi, false, true); break;            \
    default: FPDB_DEMIT(KM, kJtMark, true, true); break;                          \
  }
      switch (km) {
        case 1: FPDB_DEMIT_JT(1); break;
        case 2: FPDB_DEMIT_JT(2); break;
        default: FPDB_DEMIT_JT(0); break;
      }
#undef FPDB_DEMIT_JT
#undef FPDB_DEMIT
    }
  }
  release();   // stream-ordered reuse: later work on the stream runs after these kernels
  FPDB_HIP(hipGetLastError());
  if (d_num_out) return FPDB_OK;
  FPDB_TRY(read_count(c, num_out));
  if (*num_out > out_cap)
    return fail(FPDB_ERR_CAPACITY, "hash join probe: " + std::to_string(*num_out) +
                                       " output rows exceed the capacity of " + std::to_string(out_cap));
  return FPDB_OK;
}

static fpdb_status probe_impl(Context* c, fpdb_hash_table* t, int join_type, const fpdb_column* probe_keys,
                              int num_keys, int64_t n, const fpdb_column* probe_out, int num_probe_out,
                              const fpdb_column* build_out, int num_build_out, fpdb_column* out,
                              int64_t out_cap, int64_t* num_out, int64_t* d_num_out = nullptr) {
  KeySpec pk;
  FPDB_TRY(make_keyspec(probe_keys, num_keys, n, &pk, "hash join probe"));
  FPDB_TRY(check_key_compat(t, probe_keys, num_keys));
  int jt;
  bool track = false;
  switch (join_type) {
    case FPDB_JOIN_INNER: jt = kJtInner; break;
    case FPDB_JOIN_RIGHT: jt = kJtOuter; break;
    case FPDB_JOIN_RIGHT_SEMI: jt = kJtSemi; break;
    case FPDB_JOIN_LEFT: jt = kJtInner; track = true; break;
    case FPDB_JOIN_FULL: jt = kJtOuter; track = true; break;
    case FPDB_JOIN_LEFT_SEMI: jt = kJtMark; track = true; break;
    default: return fail(FPDB_ERR_INVALID, "unknown join type " + std::to_string(join_type));
  }
  FPDB_CHECK(!track || t->matched, FPDB_ERR_INVALID,
             "LEFT, FULL and LEFT_SEMI joins need a table built with track_build_matches");
  if (jt == kJtSemi || jt == kJtMark) num_build_out = 0;
  if (jt == kJtMark) num_probe_out = 0;
  ColSet pc, bc;
  OutColSet oc;
  FPDB_TRY(make_colset(probe_out, num_probe_out, n, &pc, "hash join probe columns"));
  FPDB_CHECK(num_build_out >= 0 && num_build_out + num_probe_out <= kMaxCols, FPDB_ERR_INVALID,
             "hash join: too many output columns");
  std::memset(&bc, 0, sizeof(bc));
  bc.n = num_build_out;
  uint32_t keymask = 0;
  for (int i = 0; i < num_build_out; ++i) {
    int w = type_width(build_out[i].type);
    FPDB_CHECK(w > 0, FPDB_ERR_INVALID, "hash join: unsupported build column type");
    // the single-column build key itself: a matching row's value is the probe key
    if (t->num_keys == 1 && build_out[i].values && build_out[i].values == t->key_values &&
        build_out[i].offset == t->key_offset && build_out[i].type == t->key_types[0])
      keymask |= 1u << i;
    if (build_out[i].values == nullptr) {
      FPDB_CHECK(t->has_payload && w == 4, FPDB_ERR_INVALID,
                 "hash join: build column without values but the table has no 4-byte inline payload");
    } else {
      FPDB_CHECK(build_out[i].length >= t->build_rows, FPDB_ERR_INVALID, "hash join: build column too short");
    }
    bc.values[i] = static_cast<const uint8_t*>(build_out[i].values);
    bc.validity[i] = build_out[i].validity;
    bc.offset[i] = build_out[i].offset;
    bc.width[i] = w;
    bc.type[i] = build_out[i].type;
  }
  int nout = (jt == kJtMark) ? 0 : num_probe_out + num_build_out;
  FPDB_TRY(make_outcolset(out, nout, &oc, "hash join output"));
  for (int i = 0; i < num_probe_out; ++i)
    FPDB_CHECK(oc.width[i] == pc.width[i], FPDB_ERR_INVALID, "hash join: output probe column width differs");
  for (int i = 0; i < num_build_out; ++i) {
    FPDB_CHECK(oc.width[num_probe_out + i] == bc.width[i], FPDB_ERR_INVALID,
               "hash join: output build column width differs");
    FPDB_CHECK(jt != kJtOuter || oc.validity[num_probe_out + i], FPDB_ERR_INVALID,
               "hash join: outer join build columns need an output validity bitmap");
    FPDB_CHECK(bc.validity[i] == nullptr || oc.validity[num_probe_out + i], FPDB_ERR_INVALID,
               "hash join: nullable build column needs an output validity bitmap");
  }
  FPDB_CHECK(out_cap >= 0, FPDB_ERR_INVALID, "negative output capacity");
  {
    // zero output validity for the rows that may be written
    ColSet none;
    std::memset(&none, 0, sizeof(none));
    for (int i = 0; i < oc.n; ++i) {
      if (!oc.validity[i]) {
        FPDB_CHECK(!(i < num_probe_out && pc.validity[i]), FPDB_ERR_INVALID,
                   "hash join: nullable probe column needs an output validity bitmap");
        continue;
      }
      FPDB_CHECK((oc.offset[i] & 31) == 0, FPDB_ERR_INVALID, "output validity offset must be a multiple of 32");
      FPDB_TRY(launch_fill(c, oc.validity[i] + oc.offset[i] / 8, 0, (uint64_t)((out_cap + 31) / 32) * 4));
    }
  }
  if (n == 0) {
    if (d_num_out) return launch_fill(c, d_num_out, 0, 8);
    *num_out = 0;
    return FPDB_OK;
  }
  if (t->build_rows == 0) {
    // nothing can match: the kernel reads no build column (an empty one may
    // have no storage); unmatched rows get nulls
    for (int i = 0; i < bc.n; ++i) bc.values[i] = nullptr;
    keymask = 0;
  }
  if (t->dense) return dense_probe(c, t, jt, track, pk, n, pc, bc, oc, out_cap, keymask, num_out, d_num_out);
  int64_t tiles = ceil_div(n, kProbeTileRows);
  FPDB_CHECK(tiles < (int64_t(1) << 31), FPDB_ERR_INVALID, "hash join probe: too many rows");
  int64_t* total = d_num_out ? d_num_out : c->d_counter;
  uint64_t mask = (uint64_t)t->capacity - 1;
  int km = key_mode_of(pk);
  uint64_t* mrow = nullptr;
  uint32_t* mcnt = nullptr;
  uint32_t* tcnt = nullptr;
  uint64_t* tbase = nullptr;
  fpdb_status st = ctx_alloc(c, (uint64_t)n * 8, reinterpret_cast<void**>(&mrow));
  if (st == FPDB_OK) st = ctx_alloc(c, (uint64_t)n * 4, reinterpret_cast<void**>(&mcnt));
  if (st == FPDB_OK) st = ctx_alloc(c, (uint64_t)tiles * 4, reinterpret_cast<void**>(&tcnt));
  if (st == FPDB_OK) st = ctx_alloc(c, (uint64_t)tiles * 8, reinterpret_cast<void**>(&tbase));
  auto release = [&]() {
    if (mrow) ctx_free(c, mrow);
    if (mcnt) ctx_free(c, mcnt);
    if (tcnt) ctx_free(c, tcnt);
    if (tbase) ctx_free(c, tbase);
  };
  if (st != FPDB_OK) {
    release();
    return st;
  }
  const int* dup = reinterpret_cast<const int*>(t->meta + 1);
  dim3 g((unsigned)tiles), b(kProbeThreads);
  {
    ProfScope ps(c, "hj_probe");
#define FPDB_MATCH_LAUNCH(KM, JT, TR)                                                                 \
  hipLaunchKernelGGL((hj_match_kernel<KM, JT, TR>), g, b, 0, c->stream, t->slots, mask, pk, n, t->matched, dup, \
                     mrow, mcnt, tcnt)
#define FPDB_MATCH_JT(KM)                                                    \
  switch (jt * 2 + (track ? 1 : 0)) {                                        \
    case kJtInner * 2: FPDB_MATCH_LAUNCH(KM, kJtInner, false); break;        \
    case kJtInner * 2 + 1: FPDB_MATCH_LAUNCH(KM, kJtInner, true); break;     \
    case kJtOuter * 2: FPDB_MATCH_LAUNCH(KM, kJtOuter, false); break;        \
    case kJtOuter * 2 + 1: FPDB_MATCH_LAUNCH(KM, kJtOuter, true); break;     \
    case kJtSemi * 2: FPDB_MATCH_LAUNCH(KM, kJtSemi, false); break;          \
    default: FPDB_MATCH_LAUNCH(KM, kJtMark, true); break;                    \
  }
    switch (km) {
      case 1: FPDB_MATCH_JT(1); break;
      case 2: FPDB_MATCH_JT(2); break;
      default: FPDB_MATCH_JT(0); break;
    }
#undef FPDB_MATCH_JT
#undef FPDB_MATCH_LAUNCH
    launch_tile_scan(c, 1, tcnt, tiles, tbase, total);
    if (jt != kJtMark && oc.n > 0 && out_cap > 0) {
#define FPDB_EMIT_LAUNCH(KM, JT)                                                                      \
  hipLaunchKernelGGL((hj_emit_kernel<KM, JT>), g, b, 0, c->stream, t->slots, mask, pk, n, pc, bc, oc, out_cap, \
                     keymask, mrow, mcnt, tbase)
#define FPDB_EMIT_JT(KM)                                           \
  switch (jt) {                                                    \
    case kJtInner: FPDB_EMIT_LAUNCH(KM, kJtInner); break;          \
    case kJtOuter: FPDB_EMIT_LAUNCH(KM, kJtOuter); break;          \
    default: FPDB_EMIT_LAUNCH(KM, kJtSemi); break;                 \
  }
      switch (km) {
        case 1: FPDB_EMIT_JT(1); break;
        case 2: FPDB_EMIT_JT(2); break;
        default: FPDB_EMIT_JT(0); break;
      }
#undef FPDB_EMIT_JT
#undef FPDB_EMIT_LAUNCH
    }
  }
  release();   // stream-ordered reuse: later work on the stream runs after these kernels
  FPDB_HIP(hipGetLastError());
  if (d_num_out) return FPDB_OK;
  FPDB_TRY(read_count(c, num_out));
  if (*num_out > out_cap)
    return fail(FPDB_ERR_CAPACITY, "hash join probe: " + std::to_string(*num_out) +
                                       " output rows exceed the capacity of " + std::to_string(out_cap));
  return FPDB_OK;
}

// Tries the dense form of a single-column table: min/max, fill, scatter, the
// duplicate check and the presence bitmap all run on the device into arrays
// sized for a span of up to twice the build rows; one read-back of
// {kmin, kmax, count, dup} decides.  Leaves t->dense null (nothing allocated)
// when the keys are absent, not distinct or too spread out.
static fpdb_status try_dense_build(Context* c, fpdb_hash_table* t, const KeySpec& k, int64_t n,
                                   const uint8_t* pv, int64_t poff) {
  const uint64_t cap = (uint64_t)std::max<int64_t>(2 * n, 1024);
  uint32_t* arr = nullptr;      // rows [cap], then payloads [cap]
  uint2* pairs = nullptr;       // {row, payload} scatter target (with a payload)
  uint64_t* bits = nullptr;
  DenseMeta* meta = nullptr;
  FPDB_TRY(ctx_alloc(c, cap * (pv ? 8 : 4), reinterpret_cast<void**>(&arr)));
  fpdb_status st = ctx_alloc(c, ((cap + 63) / 64) * 8, reinterpret_cast<void**>(&bits));
  if (st == FPDB_OK) st = ctx_alloc(c, sizeof(DenseMeta), reinterpret_cast<void**>(&meta));
  if (st == FPDB_OK && pv) st = ctx_alloc(c, cap * 8, reinterpret_cast<void**>(&pairs));
  auto drop = [&]() {
    if (arr) ctx_free(c, arr);
    if (pairs) ctx_free(c, pairs);
    if (bits) ctx_free(c, bits);
    if (meta) ctx_free(c, meta);
  };
  if (st != FPDB_OK) {
    drop();
    return st;
  }
  uint32_t* pay = pv ? arr + cap : nullptr;
  // one minmax workgroup per CU at most, so the scatter and finish
  // workgroups each reduce at most num_cus partials
  const int gm = grid_for(c, std::max<int64_t>(n, (int64_t)cap), kMinMaxThreads, 1);
  const int g = grid_for(c, n, 256, 4);
  const int gf = grid_for(c, (int64_t)cap, 256, 4);
  const int gfin = std::max(g, gf);
  // 3 words per minmax workgroup, then one filled-slot count per finish workgroup
  int64_t* parts = reinterpret_cast<int64_t*>(c->d_partials);
  FPDB_CHECK(3 * gm + gfin <= kMaxPartials, FPDB_ERR_INTERNAL, "hash join dense build: too many workgroups");
  {
    ProfScope ps(c, "hj_build");
    switch (key_mode_of(k)) {
#define FPDB_DENSE_BUILD(KM)                                                                                   \
  hipLaunchKernelGGL(hj_minmax_kernel<KM>, dim3(gm), dim3(kMinMaxThreads), 0, c->stream, k, n, parts, arr,     \
                     pairs, cap, meta);                                                                        \
  hipLaunchKernelGGL(hj_dense_scatter_kernel<KM>, dim3(g), dim3(256), 0, c->stream, arr, pairs, cap, parts, gm, \
                     k, n, pv, poff);                                                                          \
  hipLaunchKernelGGL(hj_dense_finish_kernel<KM>, dim3(gfin), dim3(256), 0, c->stream, pairs, cap, parts, gm,   \
                     meta, arr, pay, bits, parts + 3 * gm);                                                    \
  hipLaunchKernelGGL(hj_dense_decide_kernel, dim3(1), dim3(1024), 0, c->stream, parts + 3 * gm, gfin, meta)
      case 1: FPDB_DENSE_BUILD(1); break;
      case 2: FPDB_DENSE_BUILD(2); break;
      default: FPDB_DENSE_BUILD(0); break;
#undef FPDB_DENSE_BUILD
    }
  }
  DenseMeta h;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    drop();
    return fail(FPDB_ERR_HIP, std::string("hash join dense build: ") + hipGetErrorString(e));
  }
  static_assert(sizeof(DenseMeta) == 32, "four words");
  uint64_t words[4];
  st = mail_read(c, meta, 4, words);
  ctx_free(c, meta);
  meta = nullptr;
  if (pairs) ctx_free(c, pairs);   // stream-ordered: the split kernel has run
  pairs = nullptr;
  if (st != FPDB_OK) {
    drop();
    return st;
  }
  std::memcpy(&h, words, sizeof(h));
  const uint64_t span = h.count > 0 ? (uint64_t)h.kmax - (uint64_t)h.kmin + 1 : 0;
  if (h.count == 0 || span == 0 || span > cap || h.dup) {
    drop();
    return FPDB_OK;
  }
  t->dense = arr;
  t->dense_pay = pay;
  t->dense_bits = bits;
  t->dense_kmin = h.kmin;
  t->dense_span = span;
  t->capacity = (int64_t)span;
  t->inserted = h.count;
  t->unique = true;
  t->info_ready = true;
  return FPDB_OK;
}

}  // namespace fpdb

using namespace fpdb;

extern "C" {

fpdb_status fpdb_hj_build(fpdb_context* ctx, const fpdb_column* keys, int num_keys, int64_t num_rows,
                          const fpdb_column* payload, int track_build_matches, fpdb_hash_table** out) {
  FPDB_CHECK(ctx && out, FPDB_ERR_INVALID, "NULL argument");
  *out = nullptr;
  FPDB_CHECK(num_rows >= 0 && num_rows < (int64_t)kEmptyRow, FPDB_ERR_INVALID,
             "hash join build: at most 2^32-2 build rows");
  Context* c = as_ctx(ctx);
  KeySpec k;
  FPDB_TRY(make_keyspec(keys, num_keys, num_rows, &k, "hash join build"));
  if (payload) {
    FPDB_CHECK(type_width(payload->type) == 4, FPDB_ERR_INVALID, "inline payload must be a 4-byte column");
    FPDB_CHECK(payload->validity == nullptr, FPDB_ERR_INVALID, "inline payload must not have nulls");
    FPDB_CHECK(payload->length >= num_rows && (num_rows == 0 || payload->values), FPDB_ERR_INVALID,
               "inline payload column too short");
  }
  int64_t cap = 64;
  while (cap < 2 * num_rows) cap <<= 1;
  auto* t = new fpdb_hash_table();
  t->ctx = c;
  t->capacity = cap;
  t->build_rows = num_rows;
  t->num_keys = num_keys;
  for (int i = 0; i < num_keys; ++i) t->key_types[i] = keys[i].type;
  t->key_values = keys[0].values;
  t->key_offset = keys[0].offset;
  t->has_payload = payload != nullptr;
  t->payload_type = payload ? payload->type : 0;
  fpdb_status s = FPDB_OK;
  if (track_build_matches) {
    s = ctx_alloc(c, (uint64_t)std::max<int64_t>(num_rows, 1), reinterpret_cast<void**>(&t->matched));
    if (s == FPDB_OK) s = launch_fill(c, t->matched, 0, (uint64_t)std::max<int64_t>(num_rows, 1));
    if (s != FPDB_OK) { fpdb_hj_table_destroy(t); return s; }
  }
  const uint8_t* pv = payload ? static_cast<const uint8_t*>(payload->values) : nullptr;
  int64_t poff = payload ? payload->offset : 0;
  if (num_keys == 1 && num_rows > 0) {
    s = try_dense_build(c, t, k, num_rows, pv, poff);
    if (s != FPDB_OK) { fpdb_hj_table_destroy(t); return s; }
    if (t->dense) {
      *out = t;
      return FPDB_OK;
    }
  }
  s = ctx_alloc(c, (uint64_t)cap * sizeof(Slot), reinterpret_cast<void**>(&t->slots));
  if (s != FPDB_OK) { fpdb_hj_table_destroy(t); return s; }
  s = ctx_alloc(c, 32, reinterpret_cast<void**>(&t->meta));
  if (s == FPDB_OK) s = launch_fill(c, t->meta, 0, 32);
  if (s == FPDB_OK) s = launch_fill(c, t->slots, 0xFF, (uint64_t)cap * sizeof(Slot));
  if (s != FPDB_OK) { fpdb_hj_table_destroy(t); return s; }
  // meta: [0] insertion failure flag, [1] duplicate-key flag, [2] inserted rows
  int* d_err = reinterpret_cast<int*>(t->meta);
  auto* d_ins = c->d_partials;
  if (num_rows > 0) {
    dim3 g(grid_for(c, num_rows, 256)), b(256);
    ProfScope ps(c, "hj_build");
    switch (key_mode_of(k)) {
      case 1: hipLaunchKernelGGL(hj_build_kernel<1>, g, b, 0, c->stream, t->slots, (uint64_t)cap - 1, k, num_rows, pv, poff, d_err, d_ins); break;
      case 2: hipLaunchKernelGGL(hj_build_kernel<2>, g, b, 0, c->stream, t->slots, (uint64_t)cap - 1, k, num_rows, pv, poff, d_err, d_ins); break;
      default: hipLaunchKernelGGL(hj_build_kernel<0>, g, b, 0, c->stream, t->slots, (uint64_t)cap - 1, k, num_rows, pv, poff, d_err, d_ins); break;
    }
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) { fpdb_hj_table_destroy(t); return fail(FPDB_ERR_HIP, hipGetErrorString(le)); }
    s = launch_sum_partials(c, (int)g.x, t->meta + 2);
    if (s != FPDB_OK) { fpdb_hj_table_destroy(t); return s; }
    // the duplicate-key flag, read by every probe (a table of distinct keys
    // lets a probe stop at its first match)
    hipLaunchKernelGGL(hj_unique_kernel, dim3(grid_for(c, cap, 256)), dim3(256), 0, c->stream, t->slots,
                       (uint64_t)cap - 1, reinterpret_cast<int*>(t->meta + 1));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { fpdb_hj_table_destroy(t); return fail(FPDB_ERR_HIP, hipGetErrorString(e)); }
  }
  *out = t;
  return FPDB_OK;
}

fpdb_status fpdb_hj_table_destroy(fpdb_hash_table* t) {
  if (!t) return FPDB_OK;
  if (t->slots) ctx_free(t->ctx, t->slots);
  if (t->dense) ctx_free(t->ctx, t->dense);
  if (t->dense_bits) ctx_free(t->ctx, t->dense_bits);
  if (t->matched) ctx_free(t->ctx, t->matched);
  if (t->meta) ctx_free(t->ctx, t->meta);
  delete t;
  return FPDB_OK;
}

// The build is asynchronous (its duplicate-key check included); the first
// info query reads the table's counters back.
fpdb_status fpdb_hj_table_info(fpdb_hash_table* t, int64_t* build_rows, int64_t* inserted,
                               int64_t* capacity, int* unique_keys) {
  FPDB_CHECK(t, FPDB_ERR_INVALID, "NULL table");
  if (!t->info_ready) {
    Context* c = t->ctx;
    int64_t meta[3];
    FPDB_HIP(hipMemcpyAsync(meta, t->meta, sizeof(meta), hipMemcpyDeviceToHost, c->stream));
    FPDB_HIP(hipStreamSynchronize(c->stream));
    FPDB_CHECK((int)meta[0] == 0, FPDB_ERR_INTERNAL, "hash join build: table full");
    t->unique = (int)meta[1] == 0;
    t->inserted = meta[2];
    t->info_ready = true;
  }
  if (build_rows) *build_rows = t->build_rows;
  if (inserted) *inserted = t->inserted;
  if (capacity) *capacity = t->capacity;
  if (unique_keys) *unique_keys = t->unique ? 1 : 0;
  return FPDB_OK;
}

fpdb_status fpdb_hj_table_form(fpdb_hash_table* t, int* form) {
  FPDB_CHECK(t && form, FPDB_ERR_INVALID, "NULL argument");
  *form = t->dense ? 1 : 0;
  return FPDB_OK;
}

fpdb_status fpdb_hj_probe_async(fpdb_context* ctx, fpdb_hash_table* table, int join_type,
                                const fpdb_column* probe_keys, int num_keys, int64_t num_rows,
                                const fpdb_column* probe_out, int num_probe_out,
                                const fpdb_column* build_out, int num_build_out,
                                fpdb_column* out, int64_t out_capacity, int64_t* d_num_out) {
  FPDB_CHECK(ctx && table && d_num_out, FPDB_ERR_INVALID, "NULL argument");
  return probe_impl(as_ctx(ctx), table, join_type, probe_keys, num_keys, num_rows, probe_out, num_probe_out,
                    build_out, num_build_out, out, out_capacity, nullptr, d_num_out);
}

fpdb_status fpdb_hj_probe(fpdb_context* ctx, fpdb_hash_table* table, int join_type,
                          const fpdb_column* probe_keys, int num_keys, int64_t num_rows,
                          const fpdb_column* probe_out, int num_probe_out,
                          const fpdb_column* build_out, int num_build_out,
                          fpdb_column* out, int64_t out_capacity, int64_t* num_out) {
  FPDB_CHECK(ctx && table && num_out, FPDB_ERR_INVALID, "NULL argument");
  return probe_impl(as_ctx(ctx), table, join_type, probe_keys, num_keys, num_rows, probe_out,
                    num_probe_out, build_out, num_build_out, out, out_capacity, num_out);
}

fpdb_status fpdb_hj_probe_count(fpdb_context* ctx, fpdb_hash_table* table, int join_type,
                                const fpdb_column* probe_keys, int num_keys, int64_t num_rows,
                                int64_t* count) {
  FPDB_CHECK(ctx && table && count, FPDB_ERR_INVALID, "NULL argument");
  if (join_type == FPDB_JOIN_LEFT) join_type = FPDB_JOIN_INNER;
  if (join_type == FPDB_JOIN_FULL) join_type = FPDB_JOIN_RIGHT;
  if (join_type == FPDB_JOIN_LEFT_SEMI) { *count = 0; return FPDB_OK; }
  fpdb_status s = probe_impl(as_ctx(ctx), table, join_type, probe_keys, num_keys, num_rows, nullptr, 0,
                             nullptr, 0, nullptr, 0, count);
  if (s == FPDB_ERR_CAPACITY) return FPDB_OK;
  return s;
}

fpdb_status fpdb_hj_build_side_rows(fpdb_context* ctx, fpdb_hash_table* t, int matched,
                                    int64_t* d_rows, int64_t capacity, int64_t* num_out) {
  FPDB_CHECK(ctx && t && num_out, FPDB_ERR_INVALID, "NULL argument");
  FPDB_CHECK(t->matched, FPDB_ERR_INVALID, "table was built without track_build_matches");
  FPDB_CHECK(capacity >= t->build_rows && (d_rows || t->build_rows == 0), FPDB_ERR_INVALID,
             "build side rows: capacity must be at least the number of build rows");
  Context* c = as_ctx(ctx);
  ColSet none;
  OutColSet onone;
  std::memset(&none, 0, sizeof(none));
  std::memset(&onone, 0, sizeof(onone));
  onone.capacity = INT64_MAX;
  return run_compact<PredByteFlag, false>(c, PredByteFlag{t->matched, (uint8_t)(matched ? 1 : 0)}, none,
                                          onone, t->build_rows, nullptr, d_rows, num_out, nullptr,
                                          "build side rows", capacity);
}

fpdb_status fpdb_copy_rows(fpdb_context* ctx, const fpdb_column* in_all, int num_all, int64_t num_rows,
                           fpdb_column* out_all, int64_t out_row) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  FPDB_CHECK(out_row >= 0, FPDB_ERR_INVALID, "copy rows: negative output row");
  Context* c = as_ctx(ctx);
  std::vector<fpdb_column> fin, fout;
  std::vector<int> utf8;
  FPDB_TRY(split_utf8(in_all, out_all, num_all, &fin, &fout, &utf8, "copy rows"));
  for (int i : utf8) {
    FPDB_CHECK(in_all[i].length >= num_rows && out_all[i].length >= out_row + num_rows, FPDB_ERR_INVALID,
               "copy rows: column too short");
    FPDB_TRY(utf8_copy_rows(c, in_all[i], num_rows, out_all[i], out_row));
  }
  const fpdb_column* in = fin.data();
  fpdb_column* out = fout.data();
  const int num_columns = (int)fin.size();
  ColSet ic;
  OutColSet oc;
  FPDB_TRY(make_colset(in, num_columns, num_rows, &ic, "copy rows"));
  FPDB_TRY(make_outcolset(out, num_columns, &oc, "copy rows"));
  for (int i = 0; i < num_columns; ++i) {
    FPDB_CHECK(ic.width[i] == oc.width[i], FPDB_ERR_INVALID, "copy rows: output width differs");
    FPDB_CHECK(out[i].length >= out_row + num_rows, FPDB_ERR_INVALID, "copy rows: output column too short");
    FPDB_CHECK(oc.validity[i] || !ic.validity[i], FPDB_ERR_INVALID,
               "copy rows: nullable input needs an output validity bitmap");
  }
  if (num_rows == 0 || num_columns == 0) return FPDB_OK;
  hipLaunchKernelGGL(copy_rows_kernel, dim3(grid_for(c, num_rows, 256)), dim3(256), 0, c->stream, ic, num_rows,
                     oc, out_row);
  FPDB_HIP(hipGetLastError());
  return FPDB_OK;
}

fpdb_status fpdb_gather(fpdb_context* ctx, const fpdb_column* in_all, int num_all, const int64_t* d_rows,
                        int64_t num_rows, fpdb_column* out_all) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  FPDB_CHECK(d_rows || num_rows == 0, FPDB_ERR_INVALID, "NULL rows");
  Context* c = as_ctx(ctx);
  std::vector<fpdb_column> fin, fout;
  std::vector<int> utf8;
  FPDB_TRY(split_utf8(in_all, out_all, num_all, &fin, &fout, &utf8, "gather"));
  for (int i : utf8) FPDB_TRY(utf8_gather(c, in_all[i], d_rows, num_rows, out_all[i]));
  const fpdb_column* in = fin.data();
  fpdb_column* out = fout.data();
  const int num_columns = (int)fin.size();
  ColSet ic;
  OutColSet oc;
  FPDB_TRY(make_colset(in, num_columns, 0, &ic, "gather"));
  FPDB_TRY(make_outcolset(out, num_columns, &oc, "gather"));
  for (int i = 0; i < num_columns; ++i)
    FPDB_CHECK(ic.width[i] == oc.width[i], FPDB_ERR_INVALID, "gather: output width differs");
  for (int i = 0; i < oc.n; ++i) {
    if (!oc.validity[i]) continue;
    FPDB_CHECK((oc.offset[i] & 31) == 0, FPDB_ERR_INVALID, "output validity offset must be a multiple of 32");
    FPDB_TRY(launch_fill(c, oc.validity[i] + oc.offset[i] / 8, 0, (uint64_t)((num_rows + 31) / 32) * 4));
  }
  if (num_rows == 0 || num_columns == 0) return FPDB_OK;
  hipLaunchKernelGGL(gather_kernel, dim3(grid_for(c, num_rows, 256)), dim3(256), 0, c->stream, ic, d_rows,
                     num_rows, oc);
  FPDB_HIP(hipGetLastError());
  return FPDB_OK;
}

}  // extern "C"
