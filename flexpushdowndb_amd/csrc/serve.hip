// Serving loop: continuous batching of star-join queries over a star schema
// resident in HBM.
//
// The fact table is scanned in chunks, one chunk per tick, round and round (a
// circular shared scan).  Queries wait in a FIFO queue and are admitted at any
// tick into one of kServeSlots scan slots while the slot, the LDS group budget
// and the HBM workspace budget allow; a query then joins the scan at the
// current chunk and leaves it once it has seen every chunk (one full pass).
// The fact foreign keys and the measure are read once per tick for all
// queries in the scan.  Each dimension row carries one byte of query bits (bit
// q: the query in slot q keeps this row), set when a query is admitted, so a
// fact row costs one lookup per dimension for all queries together.  Rows that
// pass are aggregated per query into direct-indexed LDS group tables (count,
// sum, first row); every workgroup adds them to its own partials of that query
// (sum, count, first row + 1), so no two workgroups ever write one partial and
// no atomics reach global memory.  When a query has seen every chunk, a
// finalize workgroup adds its partials up in workgroup order and sorts the
// groups by first appearance on a second stream, writing them straight into
// pinned host memory and then a ready word; the scan stream goes on with the
// next tick meanwhile.
//
// Integer aggregation is exact and the first row of a group is a minimum, so
// a query's result does not depend on the chunk it started at or on the other
// queries: alone and in any mix it is the same.
#include "fpdb_internal.hpp"
#include "compact.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace fpdb {

constexpr int kServeSlots = 8;             // query bits per dimension row (one byte)
constexpr int kServeMaxDims = 8;
constexpr int kServeMaxFilters = 4;
constexpr int kServeGroupBudget = 8192;    // groups of all queries in one scan (16 B of LDS each)
constexpr int kServeMaxGroups = 4096;      // groups of one query (sorted in LDS at finalize)
constexpr int kServeThreads = 1024;
constexpr int kServeDefaultChunks = 8;
constexpr int kServeMaxChunks = 1024;
constexpr int kServeAheadTicks = 2;        // ticks the host may run ahead of the scan stream

struct ServeDim {
  const int32_t* fk;        // fact foreign keys, int32 (-1: outside the dimension)
  const int32_t* gattr;     // group attribute per dimension row (may be null)
  uint8_t* qmask;           // query bits per dimension row
  int64_t rows;
  int32_t domain;           // group attribute values lie in [0, domain)
};

struct ServeSlot {
  int32_t ngroup;           // group dimensions: 0 (one group), 1 or 2
  int32_t gdim[2];
  int32_t dom1;             // domain of the second group dimension
  int32_t groups;           // number of group indices, dom0 * dom1
  int32_t lds_off;          // first LDS entry of this query
  uint64_t* partial;        // [workgroup][group] x {sum, count, first row + 1 (0: none)}
};

// One tick: fact rows [row_begin, row_end) for the queries in the scan.
struct ServeBatch {
  ServeDim dim[kServeMaxDims];
  ServeSlot slot[kServeSlots];
  const int64_t* measure;
  int64_t row_begin, row_end;
  int64_t per_block;
  uint32_t users[kServeMaxDims];   // slots that join dimension d
  int32_t ndims;
  int32_t active;                  // slots in this tick
  int32_t total_groups;
};

// The finalize of one query, writing into pinned host memory.
struct ServeFinal {
  const uint64_t* partial;
  uint64_t* totals;         // per group {sum, count, first row} (first row ~0: no row)
  int32_t groups, blocks, ngroup, dom1;
  int32_t* out_g0;
  int32_t* out_g1;
  int64_t* out_cnt;
  int64_t* out_sum;
  uint64_t* out_n;          // group count, then the ready word
  uint64_t seq;
};

// Query bit of one slot for every row of a dimension: the AND of the query's
// comparisons on that dimension (none: every row).
struct ServeFilter {
  const int32_t* attr;
  int32_t op[kServeMaxFilters];
  int64_t constant[kServeMaxFilters];
  int32_t nfilters;
};

__device__ __forceinline__ bool serve_cmp(int64_t v, int32_t op, int64_t c) {
  switch (op) {
    case FPDB_CMP_LT: return v < c;
    case FPDB_CMP_LE: return v <= c;
    case FPDB_CMP_GT: return v > c;
    case FPDB_CMP_GE: return v >= c;
    case FPDB_CMP_EQ: return v == c;
    default: return v != c;
  }
}

// One byte per row, so a row's byte is written by exactly one thread.
__global__ void serve_qmask_kernel(uint8_t* qmask, int64_t rows, int slot, ServeFilter F) {
  const uint8_t bit = (uint8_t)(1u << slot);
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += (int64_t)gridDim.x * blockDim.x) {
    bool keep = true;
    for (int f = 0; f < F.nfilters; ++f) keep = keep && serve_cmp((int64_t)F.attr[r], F.op[f], F.constant[f]);
    uint8_t m = qmask[r];
    qmask[r] = keep ? (uint8_t)(m | bit) : (uint8_t)(m & ~bit);
  }
}

// int32 copy of a fact foreign-key column; keys outside [0, rows) become -1.
__global__ void serve_fk32_kernel(const int64_t* in, int64_t n, int64_t rows, int32_t* out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = in[i];
    out[i] = (v >= 0 && v < rows) ? (int32_t)v : -1;
  }
}

// Zeroes a query's partials.
__global__ void serve_clear_kernel(uint64_t* p, int64_t words) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// Group attribute columns are read through pointers kept in LDS; as such the
// compiler cannot tell that they point to global memory and would load with
// flat instructions, whose waits also hold up LDS accesses.  This view makes
// them global loads.
typedef const int32_t __attribute__((address_space(1))) GlobalI32;

__global__ __launch_bounds__(kServeThreads) void serve_scan_kernel(ServeBatch B) {
  extern __shared__ uint64_t lds[];
  // per-query descriptors indexed by a per-lane slot number live in LDS
  __shared__ int32_t s_ngroup[kServeSlots], s_gdim0[kServeSlots], s_gdim1[kServeSlots], s_dom1[kServeSlots],
      s_off[kServeSlots];
  __shared__ const int32_t* s_gattr[kServeMaxDims];
  unsigned long long* lsum = reinterpret_cast<unsigned long long*>(lds);
  unsigned* lcnt = reinterpret_cast<unsigned*>(lds + B.total_groups);
  unsigned* lfirst = lcnt + B.total_groups;
  for (int e = threadIdx.x; e < B.total_groups; e += blockDim.x) {
    lsum[e] = 0;
    lcnt[e] = 0;
    lfirst[e] = 0xFFFFFFFFu;
  }
  if (threadIdx.x < kServeSlots) {
    const int q = threadIdx.x;
    s_ngroup[q] = B.slot[q].ngroup;
    s_gdim0[q] = B.slot[q].gdim[0];
    s_gdim1[q] = B.slot[q].gdim[1];
    s_dom1[q] = B.slot[q].dom1;
    s_off[q] = B.slot[q].lds_off;
  }
  if (threadIdx.x < kServeMaxDims) s_gattr[threadIdx.x] = B.dim[threadIdx.x].gattr;
  __syncthreads();
  const int64_t begin = B.row_begin + (int64_t)blockIdx.x * B.per_block;
  const int64_t end = begin + B.per_block < B.row_end ? begin + B.per_block : B.row_end;
  for (int64_t r = begin + threadIdx.x; r < end; r += blockDim.x) {
    uint32_t m = (uint32_t)B.active;
    int32_t fkv[kServeMaxDims];
#pragma unroll
    for (int d = 0; d < kServeMaxDims; ++d) fkv[d] = 0;
#pragma unroll
    for (int d = 0; d < kServeMaxDims; ++d) {
      if (d >= B.ndims) break;
      if (!(B.users[d] & m)) continue;      // no query still in the running joins d
      const int32_t fk = B.dim[d].fk[r];
      fkv[d] = fk;
      uint32_t bits = fk >= 0 ? (uint32_t)B.dim[d].qmask[fk] : 0u;
      m &= bits | ~B.users[d];
      if (!m) break;
    }
    if (!m) continue;
    const unsigned long long meas = (unsigned long long)B.measure[r];
    const unsigned off = (unsigned)(r - begin);
    while (m) {
      const int q = __builtin_ctz(m);
      m &= m - 1;
      const int ng = s_ngroup[q];
      int g = 0;
      if (ng >= 1) {
        const int d0 = s_gdim0[q];
        int32_t f0 = 0;
        for (int d = 0; d < kServeMaxDims; ++d) if (d == d0) f0 = fkv[d];   // fkv stays in registers
        g = ((GlobalI32*)s_gattr[d0])[f0];
      }
      if (ng == 2) {
        const int d1 = s_gdim1[q];
        int32_t f1 = 0;
        for (int d = 0; d < kServeMaxDims; ++d) if (d == d1) f1 = fkv[d];
        g = g * s_dom1[q] + ((GlobalI32*)s_gattr[d1])[f1];
      }
      const int e = s_off[q] + g;
      atomicAdd(&lsum[e], meas);
      atomicAdd(&lcnt[e], 1u);
      if (off < lfirst[e]) atomicMin(&lfirst[e], off);
    }
  }
  __syncthreads();
  // add this tick's tables to this workgroup's partials (only this workgroup
  // writes them, and ticks run one after another on one stream)
  for (int e = threadIdx.x; e < B.total_groups; e += blockDim.x) {
    if (lcnt[e] == 0) continue;
    int q = -1;
    for (int s = 0; s < kServeSlots; ++s)
      if (((B.active >> s) & 1) && e >= B.slot[s].lds_off && e < B.slot[s].lds_off + B.slot[s].groups) q = s;
    if (q < 0) continue;
    const ServeSlot& S = B.slot[q];
    uint64_t* p = S.partial + 3 * ((size_t)blockIdx.x * S.groups + (e - S.lds_off));
    p[0] += lsum[e];
    p[1] += lcnt[e];
    const uint64_t f = (uint64_t)begin + lfirst[e] + 1;
    if (p[2] == 0 || f < p[2]) p[2] = f;
  }
}

// Tells the host that the scan stream has run tick `seq` (pacing only: the
// queries' results travel through their own ready words).  A one-thread
// launch after the scan: the scan's workgroups then need no shared
// completion counter, whose increments queue on one address as they finish.
__global__ void serve_tick_kernel(uint64_t* tick_flag, uint64_t seq) {
  __threadfence_system();
  __hip_atomic_store(tick_flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// A query's partials added up per group: 64 groups per workgroup, each wave
// summing every 16th workgroup's partials, then the 16 waves combined (sums
// and counts are integers, so the order does not matter; the first row is a
// minimum).
__global__ __launch_bounds__(kServeThreads) void serve_reduce_kernel(const uint64_t* partial, int32_t groups,
                                                                     int32_t blocks, uint64_t* totals) {
  __shared__ uint64_t s_acc[kServeThreads / 64][64][3];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int g = blockIdx.x * 64 + lane;
  uint64_t sum = 0, cnt = 0, first = ~0ULL;
  if (g < groups) {
    for (int b = w; b < blocks; b += kServeThreads / 64) {
      const uint64_t* p = partial + 3 * ((size_t)b * groups + g);
      const uint64_t c = p[1];
      if (!c) continue;
      sum += p[0];
      cnt += c;
      first = p[2] - 1 < first ? p[2] - 1 : first;
    }
  }
  s_acc[w][lane][0] = sum;
  s_acc[w][lane][1] = cnt;
  s_acc[w][lane][2] = first;
  __syncthreads();
  if (w == 0 && g < groups) {
    for (int v = 1; v < kServeThreads / 64; ++v) {
      sum += s_acc[v][lane][0];
      cnt += s_acc[v][lane][1];
      first = s_acc[v][lane][2] < first ? s_acc[v][lane][2] : first;
    }
    totals[3 * g] = sum;
    totals[3 * g + 1] = cnt;
    totals[3 * g + 2] = first;
  }
}

// One workgroup per query: the group totals sorted by first appearance
// (bitonic sort of first-row << 13 | group in LDS), written to pinned host
// memory, then the group count and the ready word.
__global__ __launch_bounds__(kServeThreads) void serve_finalize_kernel(ServeFinal F) {
  __shared__ unsigned long long skey[kServeMaxGroups];
  __shared__ unsigned long long s_cnt[kServeMaxGroups];
  __shared__ unsigned long long s_sum[kServeMaxGroups];
  __shared__ unsigned s_n;
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  for (int g = threadIdx.x; g < F.groups; g += blockDim.x) {
    const unsigned long long sum = F.totals[3 * g], cnt = F.totals[3 * g + 1], first = F.totals[3 * g + 2];
    s_cnt[g] = cnt;
    s_sum[g] = sum;
    if (cnt) {
      unsigned i = atomicAdd(&s_n, 1u);
      skey[i] = (first << 13) | (unsigned long long)g;
    }
  }
  __syncthreads();
  const unsigned n = s_n;
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
          bool up = (i & k) == 0;
          if ((a > b) == up) {
            skey[i] = b;
            skey[l] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (unsigned i = threadIdx.x; i < n; i += blockDim.x) {
    const int g = (int)(skey[i] & 0x1FFF);
    const int32_t g0 = F.ngroup == 0 ? 0 : (F.ngroup == 1 ? g : g / F.dom1);
    const int32_t g1 = F.ngroup == 2 ? g % F.dom1 : 0;
    __hip_atomic_store(F.out_g0 + i, g0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(F.out_g1 + i, g1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(F.out_cnt + i, (int64_t)s_cnt[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(F.out_sum + i, (int64_t)s_sum[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_store(F.out_n, (uint64_t)n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __threadfence_system();
    __hip_atomic_store(F.out_n + 1, F.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Dense dimension keys (key[i] == i) and group attributes inside their domain.
__global__ void serve_check_kernel(const int64_t* key, const int32_t* gattr, int64_t rows, int32_t domain,
                                   int* bad) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += (int64_t)gridDim.x * blockDim.x) {
    bool ok = key[r] == r;
    if (gattr) ok = ok && gattr[r] >= 0 && gattr[r] < domain;
    if (!ok) *bad = 1;
  }
}

// Pinned, device-visible host memory for one query's result: group count,
// ready word, then g0, g1 (int32) and count, sum (int64) per group.
struct ServeHostBlock {
  uint8_t* host = nullptr;
  uint8_t* dev = nullptr;
};
constexpr uint64_t kServeHostBlockBytes = 16 + (uint64_t)kServeMaxGroups * 24;
// Ready-word values handed out to finalizes, across every server of the process.
static std::atomic<uint64_t> g_serve_seq{0};

}  // namespace fpdb

struct ServeQueryRec {
  fpdb_serve_query q;
  int groups = 1;
  int state = 0;                  // 0 queued, 1 in the scan or being finalized, 2 finished, 3 failed
  int remaining = 0;              // chunks still to scan
  uint64_t workspace = 0;
  uint64_t* partial = nullptr;
  fpdb::ServeHostBlock out;
  uint64_t seq = 0;               // ready word of its finalize
  fpdb_status error = FPDB_OK;
  std::string message;
  std::vector<int32_t> g0, g1;
  std::vector<int64_t> cnt, sum;
};

struct fpdb_serve {
  fpdb::Context* ctx = nullptr;
  fpdb_serve_config cfg;
  int ndims = 0;
  fpdb::ServeDim dims[fpdb::kServeMaxDims];
  const int32_t* filter_attr[fpdb::kServeMaxDims];
  const int64_t* measure = nullptr;
  int64_t fact_rows = 0;
  int blocks = 0;
  int chunks = 1;
  int64_t chunk_rows = 0;
  int cursor = 0;                         // next chunk to scan
  hipStream_t fin = nullptr;              // finalize stream (default priority)
  hipEvent_t scanned = nullptr;           // a tick's scan, for the finalize stream to wait on
  std::mutex mu;
  std::deque<int64_t> pending;
  std::deque<ServeQueryRec> queries;
  int64_t slot_query[fpdb::kServeSlots];
  std::vector<int64_t> inflight;          // queries being finalized
  std::vector<fpdb::ServeHostBlock> host_free, host_all;
  uint64_t* tick_flag = nullptr;          // pinned: last tick the scan stream has run
  uint64_t* tick_flag_dev = nullptr;
  int32_t* fk32[fpdb::kServeMaxDims] = {};      // int32 copies of the fact foreign keys
  uint64_t ticks = 0;
  uint64_t workspace_used = 0;
};

namespace fpdb {

static void serve_free(fpdb_serve* s) {
  Context* c = s->ctx;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (s->fin) (void)hipStreamSynchronize(s->fin);
  for (int d = 0; d < s->ndims; ++d) {
    if (s->dims[d].qmask) ctx_free(c, s->dims[d].qmask);
    if (s->fk32[d]) ctx_free(c, s->fk32[d]);
  }
  for (auto& R : s->queries)
    if (R.partial) ctx_free(c, R.partial);
  for (auto& h : s->host_all) (void)hipHostFree(h.host);
  if (s->tick_flag) (void)hipHostFree(s->tick_flag);
  if (s->scanned) (void)hipEventDestroy(s->scanned);
  if (s->fin) (void)hipStreamDestroy(s->fin);
}

static int query_groups(const fpdb_serve* s, const fpdb_serve_query& q) {
  int g = 1;
  for (int i = 0; i < q.num_group; ++i) g *= s->dims[q.group_dim[i]].domain;
  return g;
}

// Per query: one partial per (workgroup, group) and one total per group, 24 B each.
static uint64_t query_workspace(const fpdb_serve* s, int groups) {
  return (uint64_t)(s->blocks + 1) * groups * 24;
}

static fpdb_status host_block(fpdb_serve* s, ServeHostBlock* out) {
  if (!s->host_free.empty()) {
    *out = s->host_free.back();
    s->host_free.pop_back();
    return FPDB_OK;
  }
  ServeHostBlock h;
  FPDB_HIP(hipHostMalloc(reinterpret_cast<void**>(&h.host), kServeHostBlockBytes,
                         hipHostMallocCoherent | hipHostMallocMapped));
  FPDB_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&h.dev), h.host, 0));
  s->host_all.push_back(h);
  *out = h;
  return FPDB_OK;
}

// Moves the queries whose finalize has written its ready word to finished.
// The caller holds the server's lock.
static void serve_collect(fpdb_serve* s, int* completed) {
  Context* c = s->ctx;
  for (size_t k = 0; k < s->inflight.size();) {
    ServeQueryRec& R = s->queries[s->inflight[k]];
    const uint64_t* words = reinterpret_cast<const uint64_t*>(R.out.host);
    if (__atomic_load_n(words + 1, __ATOMIC_ACQUIRE) != R.seq) {
      ++k;
      continue;
    }
    const size_t n = (size_t)__atomic_load_n(words, __ATOMIC_RELAXED);
    const uint8_t* base = R.out.host + 16;
    R.g0.assign(reinterpret_cast<const int32_t*>(base), reinterpret_cast<const int32_t*>(base) + n);
    R.g1.assign(reinterpret_cast<const int32_t*>(base + kServeMaxGroups * 4),
                reinterpret_cast<const int32_t*>(base + kServeMaxGroups * 4) + n);
    R.cnt.assign(reinterpret_cast<const int64_t*>(base + kServeMaxGroups * 8),
                 reinterpret_cast<const int64_t*>(base + kServeMaxGroups * 8) + n);
    R.sum.assign(reinterpret_cast<const int64_t*>(base + kServeMaxGroups * 16),
                 reinterpret_cast<const int64_t*>(base + kServeMaxGroups * 16) + n);
    s->host_free.push_back(R.out);
    R.out = ServeHostBlock();
    ctx_free(c, R.partial);
    R.partial = nullptr;
    s->workspace_used -= R.workspace;
    R.state = 2;
    ++*completed;
    s->inflight[k] = s->inflight.back();
    s->inflight.pop_back();
  }
}

static fpdb_status serve_set_qmask(fpdb_serve* s, int slot, const fpdb_serve_query& q) {
  Context* c = s->ctx;
  // every dimension's bit: the AND of this query's comparisons on it (none: 1)
  for (int d = 0; d < s->ndims; ++d) {
    ServeFilter F;
    std::memset(&F, 0, sizeof(F));
    F.attr = s->filter_attr[d];
    for (int f = 0; f < q.num_filters; ++f) {
      if (q.filter_dim[f] != d) continue;
      F.op[F.nfilters] = q.filter_op[f];
      F.constant[F.nfilters] = q.filter_constant[f];
      ++F.nfilters;
    }
    hipLaunchKernelGGL(serve_qmask_kernel, dim3(grid_for(c, s->dims[d].rows, 256, 4)), dim3(256), 0, c->stream,
                       s->dims[d].qmask, s->dims[d].rows, slot, F);
  }
  FPDB_HIP(hipGetLastError());
  return FPDB_OK;
}

// Admits queued queries into free slots (FIFO: the first query that does not
// fit the slots, the LDS group budget or the workspace budget stops
// admission, so later small queries cannot overtake it).  A query whose
// workspace cannot be allocated fails by itself.  The caller holds the
// server's lock.
static fpdb_status serve_admit(fpdb_serve* s, int* completed) {
  Context* c = s->ctx;
  int used_groups = 0;
  for (int q = 0; q < kServeSlots; ++q)
    if (s->slot_query[q] >= 0) used_groups += s->queries[s->slot_query[q]].groups;
  while (!s->pending.empty()) {
    int free_slot = -1;
    for (int q = 0; q < s->cfg.max_active && q < kServeSlots; ++q)
      if (s->slot_query[q] < 0) {
        free_slot = q;
        break;
      }
    if (free_slot < 0) break;
    const int64_t id = s->pending.front();
    ServeQueryRec& R = s->queries[id];
    const uint64_t ws = query_workspace(s, R.groups);
    if (used_groups + R.groups > kServeGroupBudget) break;
    if (s->cfg.workspace_bytes && s->workspace_used + ws > s->cfg.workspace_bytes && s->workspace_used > 0) break;
    fpdb_status st = ctx_alloc(c, ws, reinterpret_cast<void**>(&R.partial));
    if (st == FPDB_OK) st = host_block(s, &R.out);
    if (st == FPDB_OK) {
      // a pinned block may come back with the words of an earlier query (or
      // of an earlier server's, from the same pages): clear its count and
      // ready word before this query's finalize can be waited for
      uint64_t* words = reinterpret_cast<uint64_t*>(R.out.host);
      __atomic_store_n(words, 0, __ATOMIC_RELAXED);
      __atomic_store_n(words + 1, 0, __ATOMIC_RELEASE);
    }
    if (st != FPDB_OK) {
      // this query fails; the others go on
      R.error = st;
      R.message = fpdb_last_error();
      if (R.partial) ctx_free(c, R.partial);
      R.partial = nullptr;
      R.state = 3;
      s->pending.pop_front();
      ++*completed;
      continue;
    }
    const int64_t partial_words = (int64_t)s->blocks * R.groups * 3;
    hipLaunchKernelGGL(serve_clear_kernel, dim3(grid_for(c, partial_words, 256, 4)), dim3(256), 0, c->stream,
                       R.partial, partial_words);
    FPDB_HIP(hipGetLastError());
    FPDB_TRY(serve_set_qmask(s, free_slot, R.q));
    s->workspace_used += ws;
    R.workspace = ws;
    R.remaining = s->chunks;
    R.state = 1;
    used_groups += R.groups;
    s->slot_query[free_slot] = id;
    s->pending.pop_front();
  }
  return FPDB_OK;
}

// Waits until the scan stream has run tick `seq` (the pinned tick word), with
// a stream synchronize after 50 ms of spinning.
static fpdb_status wait_tick(fpdb_serve* s, uint64_t seq) {
  const auto t0 = std::chrono::steady_clock::now();
  for (uint64_t i = 0;; ++i) {
    if (__atomic_load_n(s->tick_flag, __ATOMIC_ACQUIRE) >= seq) return FPDB_OK;
    if ((i & 4095) == 4095 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) {
      FPDB_HIP(hipStreamSynchronize(s->ctx->stream));
      FPDB_CHECK(__atomic_load_n(s->tick_flag, __ATOMIC_ACQUIRE) >= seq, FPDB_ERR_INTERNAL,
                 "serve: tick signal did not arrive after the stream completed");
      return FPDB_OK;
    }
  }
}

// One tick.  The server's lock is held while queries are collected, admitted
// and booked, never while the host waits for the device.
static fpdb_status serve_step_impl(fpdb_serve* s, int* completed) {
  Context* c = s->ctx;
  *completed = 0;
  // the host runs at most kServeAheadTicks ticks ahead of the scan stream
  if (s->ticks >= (uint64_t)kServeAheadTicks) FPDB_TRY(wait_tick(s, s->ticks + 1 - kServeAheadTicks));
  ServeBatch B;
  std::memset(&B, 0, sizeof(B));
  bool scan = false, drain = false;
  std::vector<int64_t> finishing;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    serve_collect(s, completed);
    FPDB_TRY(serve_admit(s, completed));
    B.ndims = s->ndims;
    for (int d = 0; d < s->ndims; ++d) B.dim[d] = s->dims[d];
    B.measure = s->measure;
    B.row_begin = std::min<int64_t>(s->fact_rows, (int64_t)s->cursor * s->chunk_rows);
    B.row_end = std::min<int64_t>(s->fact_rows, B.row_begin + s->chunk_rows);
    B.per_block = std::max<int64_t>(1, ceil_div(B.row_end - B.row_begin, s->blocks));
    int off = 0;
    for (int q = 0; q < kServeSlots; ++q) {
      if (s->slot_query[q] < 0) continue;
      ServeQueryRec& R = s->queries[s->slot_query[q]];
      ServeSlot& S = B.slot[q];
      S.ngroup = R.q.num_group;
      S.gdim[0] = R.q.num_group >= 1 ? R.q.group_dim[0] : 0;
      S.gdim[1] = R.q.num_group >= 2 ? R.q.group_dim[1] : 0;
      S.dom1 = R.q.num_group >= 2 ? s->dims[S.gdim[1]].domain : 1;
      S.groups = R.groups;
      S.lds_off = off;
      off += R.groups;
      S.partial = R.partial;
      B.active |= 1 << q;
      for (int f = 0; f < R.q.num_filters; ++f) B.users[R.q.filter_dim[f]] |= 1u << q;
      for (int i = 0; i < R.q.num_group; ++i) B.users[R.q.group_dim[i]] |= 1u << q;
      // the query leaves the scan after this tick if it has seen every chunk;
      // its slot (and query bit) is free for the next admission, whose mask
      // kernels run after this tick's scan on the same stream
      if (--R.remaining == 0) {
        finishing.push_back(s->slot_query[q]);
        s->slot_query[q] = -1;
      }
    }
    B.total_groups = off;
    scan = B.active != 0;
    drain = !scan && s->pending.empty() && !s->inflight.empty();
    if (scan) s->cursor = (s->cursor + 1) % s->chunks;
  }
  if (scan) {
    ++s->ticks;
    {
      ProfScope ps(c, "serve_scan");
      hipLaunchKernelGGL(serve_scan_kernel, dim3(s->blocks), dim3(kServeThreads), (size_t)B.total_groups * 16,
                         c->stream, B);
    }
    hipLaunchKernelGGL(serve_tick_kernel, dim3(1), dim3(1), 0, c->stream, s->tick_flag_dev, s->ticks);
    FPDB_HIP(hipGetLastError());

    if (!finishing.empty()) {
      FPDB_HIP(hipEventRecord(s->scanned, c->stream));
      FPDB_HIP(hipStreamWaitEvent(s->fin, s->scanned, 0));
      std::lock_guard<std::mutex> lock(s->mu);
      for (int64_t id : finishing) {
        ServeQueryRec& R = s->queries[id];
        ServeFinal F;
        F.partial = R.partial;
        F.totals = R.partial + (size_t)s->blocks * R.groups * 3;
        F.groups = R.groups;
        F.blocks = s->blocks;
        F.ngroup = R.q.num_group;
        F.dom1 = R.q.num_group >= 2 ? s->dims[R.q.group_dim[1]].domain : 1;
        uint8_t* base = R.out.dev + 16;
        F.out_g0 = reinterpret_cast<int32_t*>(base);
        F.out_g1 = reinterpret_cast<int32_t*>(base + kServeMaxGroups * 4);
        F.out_cnt = reinterpret_cast<int64_t*>(base + kServeMaxGroups * 8);
        F.out_sum = reinterpret_cast<int64_t*>(base + kServeMaxGroups * 16);
        F.out_n = reinterpret_cast<uint64_t*>(R.out.dev);
        // unique in the process, so no ready word left by another server matches
        R.seq = F.seq = g_serve_seq.fetch_add(1) + 1;
        hipLaunchKernelGGL(serve_reduce_kernel, dim3((unsigned)ceil_div(R.groups, 64)), dim3(kServeThreads), 0,
                           s->fin, F.partial, F.groups, F.blocks, F.totals);
        hipLaunchKernelGGL(serve_finalize_kernel, dim3(1), dim3(kServeThreads), 0, s->fin, F);
        FPDB_HIP(hipGetLastError());
        s->inflight.push_back(id);
      }
    }
  } else if (drain) {
    // nothing to scan: wait for the finalizes still in flight
    FPDB_HIP(hipStreamSynchronize(s->fin));
    std::lock_guard<std::mutex> lock(s->mu);
    serve_collect(s, completed);
  }
  return FPDB_OK;
}

}  // namespace fpdb

using namespace fpdb;

extern "C" {

fpdb_status fpdb_serve_create(fpdb_context* ctx, const fpdb_serve_schema* schema, const fpdb_serve_config* config,
                              fpdb_serve** out) {
  FPDB_CHECK(ctx && schema && config && out, FPDB_ERR_INVALID, "NULL argument");
  *out = nullptr;
  FPDB_CHECK(schema->num_dims >= 1 && schema->num_dims <= kServeMaxDims, FPDB_ERR_INVALID, "serve: 1..8 dimensions");
  FPDB_CHECK(config->max_active >= 1 && config->max_active <= kServeSlots, FPDB_ERR_INVALID,
             "serve: max_active must be in [1, 8]");
  FPDB_CHECK(config->chunks >= 0 && config->chunks <= kServeMaxChunks, FPDB_ERR_INVALID,
             "serve: chunks must be in [0, 1024]");
  FPDB_CHECK(schema->fact_rows >= 0 && schema->fact_rows < (int64_t(1) << 40), FPDB_ERR_INVALID,
             "serve: fact rows must be in [0, 2^40)");
  FPDB_CHECK(schema->fact_measure.type == FPDB_TYPE_INT64 && schema->fact_measure.offset == 0 &&
                 !schema->fact_measure.validity && schema->fact_measure.length >= schema->fact_rows &&
                 (schema->fact_rows == 0 || schema->fact_measure.values),
             FPDB_ERR_INVALID, "serve: non-null int64 measure column needed");
  Context* c = as_ctx(ctx);
  for (int d = 0; d < schema->num_dims; ++d) {
    const fpdb_serve_dimension& D = schema->dims[d];
    FPDB_CHECK(D.key.type == FPDB_TYPE_INT64 && D.key.offset == 0 && !D.key.validity && D.key.values &&
                   D.key.length >= 1 && D.key.length < (int64_t(1) << 31),
               FPDB_ERR_INVALID, "serve: dimension keys are non-null int64 columns of 1..2^31-1 rows");
    FPDB_CHECK(D.filter_attr.type == FPDB_TYPE_INT32 && D.filter_attr.offset == 0 && D.filter_attr.values &&
                   D.filter_attr.length >= D.key.length,
               FPDB_ERR_INVALID, "serve: int32 filter attribute needed");
    FPDB_CHECK(D.fact_fk.type == FPDB_TYPE_INT64 && D.fact_fk.offset == 0 && !D.fact_fk.validity &&
                   D.fact_fk.length >= schema->fact_rows && (schema->fact_rows == 0 || D.fact_fk.values),
               FPDB_ERR_INVALID, "serve: non-null int64 fact foreign keys needed");
    if (D.group_attr.values) {
      FPDB_CHECK(D.group_attr.type == FPDB_TYPE_INT32 && D.group_attr.offset == 0 &&
                     D.group_attr.length >= D.key.length && D.group_domain >= 1 && D.group_domain <= kServeMaxGroups,
                 FPDB_ERR_INVALID, "serve: int32 group attribute with a domain in [1, 4096] needed");
    }
  }
  auto* s = new fpdb_serve();
  s->ctx = c;
  s->cfg = *config;
  s->ndims = schema->num_dims;
  s->measure = static_cast<const int64_t*>(schema->fact_measure.values);
  s->fact_rows = schema->fact_rows;
  for (int q = 0; q < kServeSlots; ++q) s->slot_query[q] = -1;
  std::memset(s->dims, 0, sizeof(s->dims));
  s->blocks = std::max(1, c->num_cus);
  s->chunks = config->chunks ? config->chunks : kServeDefaultChunks;
  s->chunk_rows = std::max<int64_t>(1, ceil_div(std::max<int64_t>(s->fact_rows, 1), s->chunks));
  fpdb_status st = FPDB_OK;
  if (hipStreamCreateWithFlags(&s->fin, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&s->scanned, hipEventDisableTiming) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&s->tick_flag), 64, hipHostMallocCoherent | hipHostMallocMapped) !=
          hipSuccess ||
      hipHostGetDevicePointer(reinterpret_cast<void**>(&s->tick_flag_dev), s->tick_flag, 0) != hipSuccess)
    st = fail(FPDB_ERR_HIP, "serve: stream, event or pinned memory creation failed");
  if (st == FPDB_OK) *s->tick_flag = 0;
  int* d_bad = nullptr;
  if (st == FPDB_OK) st = ctx_alloc(c, 8, reinterpret_cast<void**>(&d_bad));
  if (st == FPDB_OK) st = launch_fill(c, d_bad, 0, 8);
  for (int d = 0; d < s->ndims && st == FPDB_OK; ++d) {
    const fpdb_serve_dimension& D = schema->dims[d];
    ServeDim& X = s->dims[d];
    // the scan reads 4-byte foreign keys (every dimension has < 2^31 rows)
    st = ctx_alloc(c, (uint64_t)std::max<int64_t>(s->fact_rows, 1) * 4, reinterpret_cast<void**>(&s->fk32[d]));
    if (st != FPDB_OK) break;
    if (s->fact_rows > 0) {
      hipLaunchKernelGGL(serve_fk32_kernel, dim3(grid_for(c, s->fact_rows, 256)), dim3(256), 0, c->stream,
                         static_cast<const int64_t*>(D.fact_fk.values), s->fact_rows, D.key.length, s->fk32[d]);
      if (hipGetLastError() != hipSuccess) st = fail(FPDB_ERR_HIP, "serve: key copy launch failed");
    }
    X.fk = s->fk32[d];
    X.gattr = static_cast<const int32_t*>(D.group_attr.values);
    X.rows = D.key.length;
    X.domain = D.group_attr.values ? D.group_domain : 1;
    s->filter_attr[d] = static_cast<const int32_t*>(D.filter_attr.values);
    st = ctx_alloc(c, (uint64_t)X.rows, reinterpret_cast<void**>(&X.qmask));
    if (st == FPDB_OK) st = launch_fill(c, X.qmask, 0, (uint64_t)X.rows);
    if (st == FPDB_OK) {
      hipLaunchKernelGGL(serve_check_kernel, dim3(grid_for(c, X.rows, 256, 4)), dim3(256), 0, c->stream,
                         static_cast<const int64_t*>(D.key.values), X.gattr, X.rows, X.domain, d_bad);
      if (hipGetLastError() != hipSuccess) st = fail(FPDB_ERR_HIP, "serve: check launch failed");
    }
  }
  int bad = 0;
  if (st == FPDB_OK) {
    hipError_t e = hipMemcpyAsync(&bad, d_bad, 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) st = fail(FPDB_ERR_HIP, hipGetErrorString(e));
  }
  if (d_bad) ctx_free(c, d_bad);
  if (st == FPDB_OK && bad)
    st = fail(FPDB_ERR_INVALID, "serve: dimension keys must be dense (key[i] == i) and group attributes "
                                "inside their domain");
  if (st != FPDB_OK) {
    serve_free(s);
    delete s;
    return st;
  }
  *out = s;
  return FPDB_OK;
}

fpdb_status fpdb_serve_destroy(fpdb_serve* s) {
  if (!s) return FPDB_OK;
  serve_free(s);
  delete s;
  return FPDB_OK;
}

fpdb_status fpdb_serve_submit(fpdb_serve* s, const fpdb_serve_query* q, int64_t* query_id) {
  FPDB_CHECK(s && q && query_id, FPDB_ERR_INVALID, "NULL argument");
  FPDB_CHECK(q->num_filters >= 0 && q->num_filters <= kServeMaxFilters, FPDB_ERR_INVALID, "serve: 0..4 filters");
  FPDB_CHECK(q->num_group >= 0 && q->num_group <= 2, FPDB_ERR_INVALID, "serve: 0..2 group dimensions");
  for (int f = 0; f < q->num_filters; ++f) {
    FPDB_CHECK(q->filter_dim[f] >= 0 && q->filter_dim[f] < s->ndims, FPDB_ERR_INVALID, "serve: bad filter dimension");
    FPDB_CHECK(q->filter_op[f] >= FPDB_CMP_LT && q->filter_op[f] <= FPDB_CMP_NE, FPDB_ERR_INVALID,
               "serve: bad filter op");
  }
  for (int i = 0; i < q->num_group; ++i) {
    FPDB_CHECK(q->group_dim[i] >= 0 && q->group_dim[i] < s->ndims && s->dims[q->group_dim[i]].gattr,
               FPDB_ERR_INVALID, "serve: group dimensions need a group attribute");
  }
  FPDB_CHECK(q->num_group < 2 || q->group_dim[0] != q->group_dim[1], FPDB_ERR_INVALID,
             "serve: two distinct group dimensions");
  int groups = query_groups(s, *q);
  FPDB_CHECK(groups <= kServeMaxGroups, FPDB_ERR_INVALID, "serve: at most 4096 group values per query");
  FPDB_CHECK(!s->cfg.workspace_bytes || query_workspace(s, groups) <= s->cfg.workspace_bytes, FPDB_ERR_BUDGET,
             "serve: the query's workspace exceeds the budget");
  std::lock_guard<std::mutex> lock(s->mu);
  ServeQueryRec R;
  R.q = *q;
  R.groups = groups;
  s->queries.push_back(R);
  *query_id = (int64_t)s->queries.size() - 1;
  s->pending.push_back(*query_id);
  return FPDB_OK;
}

fpdb_status fpdb_serve_step(fpdb_serve* s, int* completed) {
  FPDB_CHECK(s && completed, FPDB_ERR_INVALID, "NULL argument");
  return serve_step_impl(s, completed);
}

fpdb_status fpdb_serve_status(fpdb_serve* s, int64_t query_id, int* state) {
  FPDB_CHECK(s && state, FPDB_ERR_INVALID, "NULL argument");
  std::lock_guard<std::mutex> lock(s->mu);
  FPDB_CHECK(query_id >= 0 && query_id < (int64_t)s->queries.size(), FPDB_ERR_INVALID, "serve: unknown query");
  *state = s->queries[query_id].state;
  return FPDB_OK;
}

fpdb_status fpdb_serve_result(fpdb_serve* s, int64_t query_id, int32_t* g0, int32_t* g1, int64_t* count,
                              int64_t* sum, int64_t capacity, int64_t* num_groups) {
  FPDB_CHECK(s && num_groups, FPDB_ERR_INVALID, "NULL argument");
  std::lock_guard<std::mutex> lock(s->mu);
  FPDB_CHECK(query_id >= 0 && query_id < (int64_t)s->queries.size(), FPDB_ERR_INVALID, "serve: unknown query");
  const ServeQueryRec& R = s->queries[query_id];
  if (R.state == 3) return fail(R.error, "serve: query failed: " + R.message);
  FPDB_CHECK(R.state == 2, FPDB_ERR_INVALID, "serve: query not finished");
  int64_t n = (int64_t)R.cnt.size();
  *num_groups = n;
  FPDB_CHECK(n <= capacity, FPDB_ERR_CAPACITY, "serve: result capacity too small");
  if (n) {
    if (g0) std::memcpy(g0, R.g0.data(), n * 4);
    if (g1) std::memcpy(g1, R.g1.data(), n * 4);
    if (count) std::memcpy(count, R.cnt.data(), n * 8);
    if (sum) std::memcpy(sum, R.sum.data(), n * 8);
  }
  return FPDB_OK;
}

// Replays an arrival trace in real time: queries are submitted when their
// arrival time has passed, ticks run while work is queued, scanning or being
// finalized, and each query's latency is the moment a step finds its result
// on the host minus its arrival time.
fpdb_status fpdb_serve_run_trace(fpdb_serve* s, const fpdb_serve_query* queries, const int64_t* arrival_ns,
                                 int64_t n, int64_t* query_ids, int64_t* latency_ns, int64_t* makespan_ns) {
  FPDB_CHECK(s && (n == 0 || (queries && arrival_ns && query_ids && latency_ns)), FPDB_ERR_INVALID,
             "NULL argument");
  for (int64_t i = 1; i < n; ++i)
    FPDB_CHECK(arrival_ns[i] >= arrival_ns[i - 1], FPDB_ERR_INVALID, "serve: arrivals must not decrease");
  using clock = std::chrono::steady_clock;
  const auto t0 = clock::now();
  auto now_ns = [&]() { return (int64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(clock::now() - t0).count(); };
  int64_t next = 0, done = 0;
  std::vector<int64_t> submitted;   // trace indices submitted and not yet finished
  for (int64_t i = 0; i < n; ++i) latency_ns[i] = -1;
  while (done < n) {
    int64_t t = now_ns();
    while (next < n && arrival_ns[next] <= t) {
      FPDB_TRY(fpdb_serve_submit(s, &queries[next], &query_ids[next]));
      submitted.push_back(next);
      ++next;
    }
    bool busy = false;
    {
      std::lock_guard<std::mutex> lock(s->mu);
      busy = !s->pending.empty() || !s->inflight.empty();
      for (int q = 0; q < kServeSlots && !busy; ++q) busy = s->slot_query[q] >= 0;
    }
    if (!busy) {
      if (next < n) {
        int64_t wait = arrival_ns[next] - now_ns();
        if (wait > 200000) std::this_thread::sleep_for(std::chrono::nanoseconds(wait - 100000));
        continue;
      }
      return fail(FPDB_ERR_INTERNAL, "serve: trace stalled");
    }
    int completed = 0;
    FPDB_TRY(serve_step_impl(s, &completed));
    if (!completed) continue;
    const int64_t t_done = now_ns();
    std::lock_guard<std::mutex> lock(s->mu);
    for (size_t k = 0; k < submitted.size();) {
      const int64_t i = submitted[k];
      const int state = s->queries[query_ids[i]].state;
      if (state < 2) {
        ++k;
        continue;
      }
      latency_ns[i] = state == 2 ? t_done - arrival_ns[i] : -1;
      ++done;
      submitted[k] = submitted.back();
      submitted.pop_back();
    }
  }
  if (makespan_ns) *makespan_ns = now_ns();
  return FPDB_OK;
}

}  // extern "C"
