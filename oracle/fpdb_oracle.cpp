// fpdb_oracle.cpp -- CPU restatement of FlexPushdownDB's relational hot path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
// calls this file.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline use it, and only to check results (or as the timed CPU baseline),
// never as the thing measured on the GPU.
//
// Every function cites the reference file:line whose behaviour it restates.
// The reference paths are relative to the FlexPushdownDB source tree.
//
// Pinning status (how each function is checked before it is trusted):
//   * bloom_masks            pinned: independent numpy MT19937 + seed_seq
//                            restatement, std::mt19937 10000th-output known
//                            answer, 4..5-bits-per-window invariant.
//   * arrow6_hash_*          PARITY UNPINNED: Arrow 6.0.0 key_hash.cc is not in
//                            the reference tree; restated from its published
//                            algorithm (helper_8B).  Checked only for the
//                            closed-form properties (bijective multiply, byte
//                            swap, zero -> zero, null -> 0).
//   * blocked insert/find    pinned: no false negatives, exact FPR identity on
//                            tiny brute force, numpy restatement of block/mask.
//   * legacy bloom           pinned: k/m closed forms, primality by trial
//                            division on small ranges, hashCombine worked
//                            example, 128-bit exact arithmetic cross-check.
//   * hash join              pinned: closed form for a permutation build side,
//                            brute force nested loop, ATTIC JoinTest vector.
//   * group-by               pinned: dense direct-array closed form, ATTIC
//                            GroupTest vector, exact math.fsum for FP64.
#include <cstdint>
#include <cstring>
#include <cmath>
#include <random>
#include <vector>
#include <unordered_map>
#include <algorithm>

extern "C" {

// ---------------------------------------------------------------------------
// Blocked bloom filter (vendored Arrow BlockedBloomFilter)
// ---------------------------------------------------------------------------

static const int kBitsPerMask = 57;            // BloomFilter.h:66
static const int kMinBitsSet = 4;              // BloomFilter.h:74
static const int kMaxBitsSet = 5;              // BloomFilter.h:75
static const int kLogNumMasks = 10;            // BloomFilter.h:83
static const int kNumMasks = 1 << kLogNumMasks;
static const int kTotalBytes = (kNumMasks + 64) / 8;   // BloomFilter.h:89 (136)

static inline int get_bit(const uint8_t* v, int64_t i) { return (v[i >> 3] >> (i & 7)) & 1; }
static inline void set_bit(uint8_t* v, int64_t i) { v[i >> 3] |= (uint8_t)(1u << (i & 7)); }

// Restates BloomFilterMasks::BloomFilterMasks()
// (fpdb-tuple/src/arrow/exec/BloomFilter.cpp:16-79).  The sequence depends on
// libstdc++'s std::uniform_int_distribution<uint64_t>, so this file must be
// compiled with g++/libstdc++ (as the reference is, cmake/CMakeUtils.cmake:33-38).
int orc_bloom_masks(uint8_t* out /* 136 bytes */) {
  std::seed_seq seed{0, 0, 0, 0, 0, 0, 0, 0};
  std::mt19937 re(seed);
  std::uniform_int_distribution<uint64_t> rd;
  auto random = [&re, &rd](int lo, int hi) -> uint64_t {
    return (uint64_t)lo + rd(re) % (uint64_t)(hi - lo + 1);
  };
  std::memset(out, 0, kTotalBytes);
  int num_bits_set = (int)random(kMinBitsSet, kMaxBitsSet);
  for (int i = 0; i < num_bits_set; ++i) {
    for (;;) {
      int pos = (int)random(0, kBitsPerMask - 1);
      if (!get_bit(out, pos)) { set_bit(out, pos); break; }
    }
  }
  const int64_t total = kNumMasks + kBitsPerMask - 1;
  for (int64_t i = kBitsPerMask; i < total; ++i) {
    int leaving = get_bit(out, i - kBitsPerMask);
    if (leaving == 1 && num_bits_set == kMinBitsSet) { set_bit(out, i); continue; }
    if (leaving == 0 && num_bits_set == kMaxBitsSet) { continue; }
    if (random(0, kBitsPerMask * 2 - 1) < (uint64_t)(kMinBitsSet + kMaxBitsSet)) {
      set_bit(out, i);
      if (leaving == 0) ++num_bits_set;
    } else {
      if (leaving == 1) --num_bits_set;
    }
  }
  return kTotalBytes;
}

// First `count` raw outputs of std::mt19937 seeded with seed_seq{0 x 8} and of
// a default-constructed std::mt19937 (for the standard's known answer).
void orc_mt19937_raw(int default_seeded, uint32_t* out, int64_t count) {
  if (default_seeded) {
    std::mt19937 g;
    for (int64_t i = 0; i < count; ++i) out[i] = (uint32_t)g();
  } else {
    std::seed_seq seed{0, 0, 0, 0, 0, 0, 0, 0};
    std::mt19937 g(seed);
    for (int64_t i = 0; i < count; ++i) out[i] = (uint32_t)g();
  }
}

// uniform_int_distribution<uint64_t> full-range draws on the seeded engine.
void orc_uniform_u64_draws(uint64_t* out, int64_t count) {
  std::seed_seq seed{0, 0, 0, 0, 0, 0, 0, 0};
  std::mt19937 re(seed);
  std::uniform_int_distribution<uint64_t> rd;
  for (int64_t i = 0; i < count; ++i) out[i] = rd(re);
}

// BloomFilterMasks::mask (BloomFilter.h:52-61): unaligned little-endian 64-bit
// load at byte bit_offset/8, shifted by bit_offset%8, masked to 57 bits.
static inline uint64_t mask_at(const uint8_t* masks, int bit_offset) {
  uint64_t w;
  std::memcpy(&w, masks + bit_offset / 8, 8);
  return (w >> (bit_offset % 8)) & ((1ULL << kBitsPerMask) - 1);
}

static inline uint64_t rotl64(uint64_t x, int n) {   // Util.h:10
  return (x << n) | (x >> ((-n) & 63));
}

// BlockedBloomFilter::mask (BloomFilter.h:190-205)
static inline uint64_t bb_mask(const uint8_t* masks, uint64_t hash) {
  int id = (int)(hash & (kNumMasks - 1));
  uint64_t m = mask_at(masks, id);
  int rot = (int)((hash >> kLogNumMasks) & 63);
  return rotl64(m, rot);
}

// BlockedBloomFilter::block_id (BloomFilter.h:207-212)
static inline int64_t bb_block(uint64_t hash, int64_t num_blocks) {
  return (int64_t)((hash >> (kLogNumMasks + 6)) & (uint64_t)(num_blocks - 1));
}

// BlockedBloomFilter::CreateEmpty sizing (BloomFilter.cpp:87-99):
// log_num_bits = ceil(log2(max(512, 8 * n))) (Arrow BitUtil::Log2 is a ceiling).
int orc_blocked_log_num_blocks(int64_t num_rows_to_insert) {
  int64_t desired = std::max<int64_t>(512, num_rows_to_insert * 8);
  int log_bits = 0;
  while ((int64_t(1) << log_bits) < desired) ++log_bits;
  return log_bits - 6;
}

// BlockedBloomFilter::InsertImp / Insert (BloomFilter.cpp:109-114, BloomFilter.h:184-188)
void orc_blocked_insert(const uint8_t* masks, uint64_t* blocks, int log_num_blocks,
                        const uint32_t* hashes, int64_t n) {
  int64_t nb = int64_t(1) << log_num_blocks;
  for (int64_t i = 0; i < n; ++i) {
    uint64_t h = hashes[i];
    blocks[bb_block(h, nb)] |= bb_mask(masks, h);
  }
}

// BlockedBloomFilter::FindImp (BloomFilter.cpp:138-170): bit i of the result is
// Find(hashes[i]); full 64-bit words, then ceil((n % 64) / 8) tail bytes.
void orc_blocked_find(const uint8_t* masks, const uint64_t* blocks, int log_num_blocks,
                      const uint32_t* hashes, int64_t n, uint8_t* result) {
  int64_t nb = int64_t(1) << log_num_blocks;
  uint64_t bits = 0;
  for (int64_t i = 0; i < n; ++i) {
    uint64_t h = hashes[i];
    uint64_t m = bb_mask(masks, h);
    uint64_t b = blocks[bb_block(h, nb)];
    uint64_t r = ((b & m) == m) ? 1ULL : 0ULL;
    bits |= r << (i & 63);
    if ((i & 63) == 63) {
      std::memcpy(result + (i / 64) * 8, &bits, 8);
      bits = 0;
    }
  }
  int tail = (int)(((n % 64) + 7) / 8);
  for (int i = 0; i < tail; ++i) result[n / 64 * 8 + i] = (uint8_t)(bits >> (i * 8));
}

// ---------------------------------------------------------------------------
// Arrow 6.0.0 key hash (third-party, not in the reference tree).
// Called at fpdb-tuple/src/RecordBatchHasher.cpp:97-100 through
// arrow::compute::Hashing::HashMultiColumn.  Restated from Arrow 6.0.0
// cpp/src/arrow/compute/exec/key_hash.cc: a fixed-width key of w <= 8 bytes is
// zero-extended to 64 bits, multiplied by 0xC2B2AE3D27D4EB4F, byte-swapped and
// truncated to 32 bits (helper_8B); a null key hashes to 0; column i > 0 is
// folded in with h ^= hi + 0x9e3779b9 + (h << 6) + (h >> 2).
// PARITY UNPINNED against the reference binary.
// ---------------------------------------------------------------------------
static inline uint32_t arrow6_hash_word(uint64_t x) {
  return (uint32_t)__builtin_bswap64(x * 0xC2B2AE3D27D4EB4FULL);
}

// width = 4 or 8; validity may be NULL (Arrow LSB-first bitmap, bit offset 0)
void orc_arrow6_hash_column(const void* values, int width, const uint8_t* validity,
                            int64_t n, uint32_t* out, int combine) {
  for (int64_t i = 0; i < n; ++i) {
    uint64_t x = 0;
    if (width == 8) x = ((const uint64_t*)values)[i];
    else if (width == 4) x = ((const uint32_t*)values)[i];
    else if (width == 2) x = ((const uint16_t*)values)[i];
    else x = ((const uint8_t*)values)[i];
    uint32_t h = arrow6_hash_word(x);
    if (validity && !get_bit(validity, i)) h = 0;
    if (combine) {
      uint32_t p = out[i];
      out[i] = p ^ (h + 0x9e3779b9u + (p << 6) + (p >> 2));
    } else {
      out[i] = h;
    }
  }
}

// ---------------------------------------------------------------------------
// Legacy bloom filter (fpdb-executor/src/physical/bloomfilter/BloomFilter.cpp,
// UniversalHashFunction.cpp).
// ---------------------------------------------------------------------------

// BloomFilter::k_from_p (BloomFilter.cpp:249-252)
int64_t orc_legacy_k(double p) { return (int64_t)std::ceil(std::log(1 / p) / std::log(2)); }

// BloomFilter::m_from_np (BloomFilter.cpp:254-257)
int64_t orc_legacy_m(int64_t n, double p) {
  return (int64_t)std::ceil(((double)n) * std::abs(std::log(p)) / std::pow(std::log(2), 2));
}

// ArrayHasher::hash -> fpdb::util::hashCombine (fpdb-tuple/src/ArrayHasher.cpp:12-19,
// fpdb-util/src/Util.cpp:106-113) over std::hash<int32/int64> (identity in
// libstdc++): seed = number of columns, seed ^= h + 0x9e3779b9 + (seed<<6) + (seed>>2).
// cols: array of `ncols` pointers, widths in {4, 8} (signed values).
void orc_legacy_keys(const void* const* cols, const int* widths, int ncols, int64_t n,
                     int64_t* out) {
  for (int64_t i = 0; i < n; ++i) {
    uint64_t seed = (uint64_t)ncols;
    for (int c = 0; c < ncols; ++c) {
      uint64_t h;
      if (widths[c] == 8) h = (uint64_t)((const int64_t*)cols[c])[i];
      else h = (uint64_t)(int64_t)((const int32_t*)cols[c])[i];
      seed ^= h + 0x9e3779b9ULL + (seed << 6) + (seed >> 2);
    }
    out[i] = (int64_t)seed;
  }
}

// UniversalHashFunction::hash (UniversalHashFunction.cpp:51-57):
// sum = (uint64)(a*x + b) with int64 wraparound; h = (sum % p) % m in uint64.
static inline int64_t uhash(int64_t a, int64_t b, int64_t m, int64_t p, int64_t x) {
  uint64_t sum = (uint64_t)a * (uint64_t)x + (uint64_t)b;
  return (int64_t)((sum % (uint64_t)p) % (uint64_t)m);
}

// BloomFilter::add (BloomFilter.cpp:62-73); bit h -> word h/64, bit h%64.
// params: k rows of {a, b, m, p}.
void orc_legacy_add(const int64_t* params, int k, int64_t* words, const int64_t* keys, int64_t n) {
  for (int64_t i = 0; i < n; ++i) {
    for (int j = 0; j < k; ++j) {
      const int64_t* q = params + 4 * j;
      int64_t h = uhash(q[0], q[1], q[2], q[3], keys[i]);
      words[h / 64] |= (int64_t)(1ULL << (h % 64));
    }
  }
}

// BloomFilter::contains (BloomFilter.cpp:75-91); capacity == 0 -> false.
// Result is a packed LSB-first selection bitmap (the gandiva SelectionVector of
// BloomFilterUseKernel.cpp:159-167 holds exactly the rows whose bit is set).
void orc_legacy_find(const int64_t* params, int k, int64_t capacity, const int64_t* words,
                     const int64_t* keys, int64_t n, uint8_t* result) {
  std::memset(result, 0, (size_t)((n + 7) / 8));
  if (capacity == 0) return;
  for (int64_t i = 0; i < n; ++i) {
    bool ok = true;
    for (int j = 0; j < k && ok; ++j) {
      const int64_t* q = params + 4 * j;
      int64_t h = uhash(q[0], q[1], q[2], q[3], keys[i]);
      if (!(((uint64_t)words[h / 64] >> (h % 64)) & 1ULL)) ok = false;
    }
    if (ok) set_bit(result, i);
  }
}

// primesieve::nth_prime(0, i): the smallest prime >= i
// (UniversalHashFunction.cpp:27-28).  Deterministic Miller-Rabin for 64-bit.
static uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (unsigned __int128)a * b % m; }
static uint64_t powmod(uint64_t a, uint64_t e, uint64_t m) {
  uint64_t r = 1; a %= m;
  while (e) { if (e & 1) r = mulmod(r, a, m); a = mulmod(a, a, m); e >>= 1; }
  return r;
}
int orc_is_prime(uint64_t n) {
  if (n < 2) return 0;
  static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  for (uint64_t p : small) { if (n % p == 0) return n == p; }
  uint64_t d = n - 1; int s = 0;
  while ((d & 1) == 0) { d >>= 1; ++s; }
  for (uint64_t a : small) {
    uint64_t x = powmod(a, d, n);
    if (x == 1 || x == n - 1) continue;
    bool comp = true;
    for (int r = 1; r < s; ++r) { x = mulmod(x, x, n); if (x == n - 1) { comp = false; break; } }
    if (comp) return 0;
  }
  return 1;
}
uint64_t orc_next_prime(uint64_t i) { while (!orc_is_prime(i)) ++i; return i; }

// UniversalHashFunction(m) (UniversalHashFunction.cpp:13-35) with the
// std::random_device seed replaced by an explicit seed so the parameters are
// reproducible inputs: i ~ U[m, 2m], p = nth_prime(0, i), a ~ U[1, p-1],
// b ~ U[0, p-1], each drawn with std::uniform_int_distribution<int64_t> from
// std::mt19937.  out: k rows of {a, b, m, p}.
void orc_legacy_make_params(uint32_t seed, int64_t m, int k, int64_t* out) {
  std::mt19937 gen(seed);
  for (int j = 0; j < k; ++j) {
    int64_t i = std::uniform_int_distribution<int64_t>(m, 2 * m)(gen);
    int64_t p = (int64_t)orc_next_prime((uint64_t)i);
    int64_t a = std::uniform_int_distribution<int64_t>(1, p - 1)(gen);
    int64_t b = std::uniform_int_distribution<int64_t>(0, p - 1)(gen);
    out[4 * j + 0] = a; out[4 * j + 1] = b; out[4 * j + 2] = m; out[4 * j + 3] = p;
  }
}

// ---------------------------------------------------------------------------
// Hash join (inner / outer / semi), restating the semantics of
// HashJoinArrowKernel (fpdb-executor/src/physical/join/hashjoin/HashJoinArrowKernel.cpp:126-206)
// and the legacy TupleSetIndex multimap (fpdb-tuple/src/TupleSetIndex.cpp:44-88,
// RecordBatchHashJoiner.cpp:83-119).  Null keys never match.
// Output pairs are in probe-row order; for one probe row, matches follow
// build-row order.  Keys are int64 (int32 callers widen).
// Returns the number of pairs; if out arrays are NULL only counts.
// ---------------------------------------------------------------------------
int64_t orc_join_pairs(const int64_t* bkeys, const uint8_t* bvalid, int64_t nb,
                       const int64_t* pkeys, const uint8_t* pvalid, int64_t np,
                       int64_t* out_probe, int64_t* out_build, uint8_t* build_matched) {
  std::unordered_map<int64_t, std::vector<int64_t>> index;
  index.reserve((size_t)nb * 2 + 1);
  for (int64_t i = 0; i < nb; ++i) {
    if (bvalid && !get_bit(bvalid, i)) continue;
    index[bkeys[i]].push_back(i);
  }
  int64_t cnt = 0;
  for (int64_t r = 0; r < np; ++r) {
    if (pvalid && !get_bit(pvalid, r)) continue;
    auto it = index.find(pkeys[r]);
    if (it == index.end()) continue;
    for (int64_t b : it->second) {
      if (out_probe) { out_probe[cnt] = r; out_build[cnt] = b; }
      if (build_matched) build_matched[b] = 1;
      ++cnt;
    }
  }
  return cnt;
}

// ---------------------------------------------------------------------------
// Hash group-by (GroupArrowKernel -> Arrow "aggregate" node with hash_count /
// hash_sum, GroupArrowKernel.cpp:202-243, Sum.cpp:63-74, Count.cpp:91-104).
// Groups are numbered in order of first appearance (the serial grouper's
// group-id order).  COUNT counts rows (ALL); integer SUM wraps in int64; FP64
// SUM is accumulated sequentially in row order.
// keys: packed uint64 group key per row.
// vals_i64[j] (j < n_i64) and vals_f64[j] (j < n_f64) are aggregate inputs.
// Outputs sized for at most `max_groups` groups; returns the group count or -1.
// ---------------------------------------------------------------------------
int64_t orc_groupby(const uint64_t* keys, int64_t n,
                    const int64_t* const* vals_i64, int n_i64,
                    const double* const* vals_f64, int n_f64,
                    int64_t max_groups,
                    uint64_t* out_keys, int64_t* out_count,
                    int64_t* out_sum_i64 /* n_i64 x max_groups */,
                    double* out_sum_f64 /* n_f64 x max_groups */,
                    int64_t* out_first_row) {
  std::unordered_map<uint64_t, int64_t> gid;
  gid.reserve(1 << 16);
  int64_t ng = 0;
  for (int64_t r = 0; r < n; ++r) {
    auto it = gid.find(keys[r]);
    int64_t g;
    if (it == gid.end()) {
      if (ng >= max_groups) return -1;
      g = ng++;
      gid.emplace(keys[r], g);
      out_keys[g] = keys[r];
      out_count[g] = 0;
      out_first_row[g] = r;
      for (int j = 0; j < n_i64; ++j) out_sum_i64[(int64_t)j * max_groups + g] = 0;
      for (int j = 0; j < n_f64; ++j) out_sum_f64[(int64_t)j * max_groups + g] = 0.0;
    } else {
      g = it->second;
    }
    out_count[g] += 1;
    for (int j = 0; j < n_i64; ++j) {
      int64_t* s = &out_sum_i64[(int64_t)j * max_groups + g];
      *s = (int64_t)((uint64_t)*s + (uint64_t)vals_i64[j][r]);
    }
    for (int j = 0; j < n_f64; ++j) out_sum_f64[(int64_t)j * max_groups + g] += vals_f64[j][r];
  }
  return ng;
}

}  // extern "C"
