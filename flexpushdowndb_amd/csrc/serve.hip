// Serving loop: continuous batching of star-join queries over a star schema
// resident in HBM.
//
// Queries wait in a FIFO queue and are admitted into one of kServeSlots scan
// slots while the slot, the LDS group budget and the HBM workspace budget
// allow.  One step scans the whole fact table once for every admitted query
// (a shared scan): the fact foreign keys and the measure are read once per
// step, whatever the number of queries.  Each dimension row carries one byte
// of query bits (bit q: the query in slot q keeps this row), set when a query
// is admitted, so a fact row costs one lookup per dimension for all queries
// together.  Rows that pass are aggregated per query into direct-indexed LDS
// group tables (count, sum, first row); every workgroup writes its partials,
// and a finalize workgroup per query adds them up in a fixed order and sorts
// the groups by first appearance.  Queries that arrive during a step are
// admitted at the next one.
//
// Integer aggregation is exact and the group order depends only on the fact
// rows, so a query's result is the same alone and in any mix.
#include "fpdb_internal.hpp"
#include "compact.hpp"

#include <algorithm>
#include <chrono>
#include <climits>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

namespace fpdb {

constexpr int kServeSlots = 8;             // query bits per dimension row (one byte)
constexpr int kServeMaxDims = 8;
constexpr int kServeMaxFilters = 4;
constexpr int kServeGroupBudget = 8192;    // groups of all queries in one scan (16 B of LDS each)
constexpr int kServeMaxGroups = 4096;      // groups of one query (sorted in LDS at finalize)
constexpr int kServeThreads = 1024;

struct ServeDim {
  const int64_t* fk;        // fact foreign keys
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
  uint64_t* partial;        // [workgroup][group] x {sum, count << 32 | first row offset}
  int32_t* out_g0;
  int32_t* out_g1;
  int64_t* out_cnt;
  int64_t* out_sum;
  unsigned long long* out_n;
};

struct ServeBatch {
  ServeDim dim[kServeMaxDims];
  ServeSlot slot[kServeSlots];
  const int64_t* measure;
  int64_t fact_rows;
  int64_t per_block;
  uint32_t users[kServeMaxDims];   // slots that join dimension d
  int32_t ndims;
  int32_t active;                  // slots in this scan
  int32_t total_groups;
  int32_t blocks;
  int32_t finalize_slot[kServeSlots];
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
  const int64_t begin = (int64_t)blockIdx.x * B.per_block;
  const int64_t end = begin + B.per_block < B.fact_rows ? begin + B.per_block : B.fact_rows;
  for (int64_t r = begin + threadIdx.x; r < end; r += blockDim.x) {
    uint32_t m = (uint32_t)B.active;
    int64_t fkv[kServeMaxDims];
#pragma unroll
    for (int d = 0; d < kServeMaxDims; ++d) fkv[d] = 0;
#pragma unroll
    for (int d = 0; d < kServeMaxDims; ++d) {
      if (d >= B.ndims) break;
      if (!(B.users[d] & m)) continue;      // no query still in the running joins d
      int64_t fk = B.dim[d].fk[r];
      fkv[d] = fk;
      uint32_t bits = (fk >= 0 && fk < B.dim[d].rows) ? (uint32_t)B.dim[d].qmask[fk] : 0u;
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
        int64_t f0 = 0;
        for (int d = 0; d < kServeMaxDims; ++d) if (d == d0) f0 = fkv[d];   // fkv stays in registers
        g = s_gattr[d0][f0];
      }
      if (ng == 2) {
        const int d1 = s_gdim1[q];
        int64_t f1 = 0;
        for (int d = 0; d < kServeMaxDims; ++d) if (d == d1) f1 = fkv[d];
        g = g * s_dom1[q] + s_gattr[d1][f1];
      }
      const int e = s_off[q] + g;
      atomicAdd(&lsum[e], meas);
      atomicAdd(&lcnt[e], 1u);
      if (off < lfirst[e]) atomicMin(&lfirst[e], off);
    }
  }
  __syncthreads();
  // this workgroup's partials, one entry per (query, group)
  for (int e = threadIdx.x; e < B.total_groups; e += blockDim.x) {
    int q = -1;
    for (int s = 0; s < kServeSlots; ++s)
      if (((B.active >> s) & 1) && e >= B.slot[s].lds_off && e < B.slot[s].lds_off + B.slot[s].groups) q = s;
    if (q < 0) continue;
    const ServeSlot& S = B.slot[q];
    uint64_t* p = S.partial + 2 * ((size_t)blockIdx.x * S.groups + (e - S.lds_off));
    p[0] = lsum[e];
    p[1] = ((uint64_t)lcnt[e] << 32) | lfirst[e];
  }
}

// One workgroup per query: partials added in workgroup order, groups sorted by
// first appearance (bitonic sort of first-row << 13 | group in LDS).
__global__ __launch_bounds__(kServeThreads) void serve_finalize_kernel(ServeBatch B) {
  __shared__ unsigned long long skey[kServeMaxGroups];
  __shared__ unsigned s_n;
  const int q = B.finalize_slot[blockIdx.x];
  const ServeSlot& S = B.slot[q];
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  for (int g = threadIdx.x; g < S.groups; g += blockDim.x) {
    unsigned long long sum = 0, cnt = 0;
    unsigned long long first = ~0ULL;
    for (int b = 0; b < B.blocks; ++b) {
      const uint64_t* p = S.partial + 2 * ((size_t)b * S.groups + g);
      unsigned c = (unsigned)(p[1] >> 32);
      if (!c) continue;
      sum += p[0];
      cnt += c;
      unsigned long long f = (unsigned long long)b * B.per_block + (unsigned)(p[1] & 0xFFFFFFFFu);
      first = f < first ? f : first;
    }
    S.out_cnt[g] = (int64_t)cnt;     // by group index for now
    S.out_sum[g] = (int64_t)sum;
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
  // the sorted groups, then their counts and sums (read by group index before
  // any is overwritten in rank order)
  unsigned long long c_tmp[kServeMaxGroups / kServeThreads], s_tmp[kServeMaxGroups / kServeThreads];
  for (unsigned i = threadIdx.x, t = 0; i < n; i += blockDim.x, ++t) {
    int g = (int)(skey[i] & 0x1FFF);
    c_tmp[t] = (unsigned long long)S.out_cnt[g];
    s_tmp[t] = (unsigned long long)S.out_sum[g];
  }
  __syncthreads();
  for (unsigned i = threadIdx.x, t = 0; i < n; i += blockDim.x, ++t) {
    int g = (int)(skey[i] & 0x1FFF);
    S.out_g0[i] = S.ngroup == 0 ? 0 : (S.ngroup == 1 ? g : g / S.dom1);
    S.out_g1[i] = S.ngroup == 2 ? g % S.dom1 : 0;
    S.out_cnt[i] = (int64_t)c_tmp[t];
    S.out_sum[i] = (int64_t)s_tmp[t];
  }
  if (threadIdx.x == 0) *S.out_n = n;
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

}  // namespace fpdb

struct ServeQueryRec {
  fpdb_serve_query q;
  int groups = 1;
  int state = 0;                  // 0 queued, 1 in a scan, 2 done
  uint64_t workspace = 0;
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
  int64_t per_block = 0;
  std::mutex mu;
  std::deque<int64_t> pending;
  std::deque<ServeQueryRec> queries;
  int64_t slot_query[fpdb::kServeSlots];
  uint64_t* slot_partial[fpdb::kServeSlots];
  // per-slot result buffers (kServeMaxGroups entries)
  int32_t* res_g0[fpdb::kServeSlots];
  int32_t* res_g1[fpdb::kServeSlots];
  int64_t* res_cnt[fpdb::kServeSlots];
  int64_t* res_sum[fpdb::kServeSlots];
  unsigned long long* res_n = nullptr;   // kServeSlots counters
  uint64_t workspace_used = 0;
  int64_t steps = 0;
};

namespace fpdb {

static void serve_free(fpdb_serve* s) {
  Context* c = s->ctx;
  for (int d = 0; d < s->ndims; ++d)
    if (s->dims[d].qmask) ctx_free(c, s->dims[d].qmask);
  for (int q = 0; q < kServeSlots; ++q) {
    if (s->slot_partial[q]) ctx_free(c, s->slot_partial[q]);
    if (s->res_g0[q]) ctx_free(c, s->res_g0[q]);
    if (s->res_g1[q]) ctx_free(c, s->res_g1[q]);
    if (s->res_cnt[q]) ctx_free(c, s->res_cnt[q]);
    if (s->res_sum[q]) ctx_free(c, s->res_sum[q]);
  }
  if (s->res_n) ctx_free(c, s->res_n);
}

static int query_groups(const fpdb_serve* s, const fpdb_serve_query& q) {
  int g = 1;
  for (int i = 0; i < q.num_group; ++i) g *= s->dims[q.group_dim[i]].domain;
  return g;
}

static uint64_t query_workspace(const fpdb_serve* s, int groups) {
  return (uint64_t)s->blocks * groups * 16;
}

// Admits queued queries into free slots (FIFO: the first query that does not
// fit stops admission, so later small queries cannot overtake it).  The
// caller holds the server's lock.
static fpdb_status serve_admit(fpdb_serve* s, std::vector<int>& admitted) {
  Context* c = s->ctx;
  int used_groups = 0;
  for (int q = 0; q < kServeSlots; ++q)
    if (s->slot_query[q] >= 0) used_groups += s->queries[s->slot_query[q]].groups;
  while (!s->pending.empty()) {
    int free_slot = -1;
    for (int q = 0; q < s->cfg.max_active && q < kServeSlots; ++q)
      if (s->slot_query[q] < 0) { free_slot = q; break; }
    if (free_slot < 0) break;
    ServeQueryRec& R = s->queries[s->pending.front()];
    uint64_t ws = query_workspace(s, R.groups);
    if (used_groups + R.groups > kServeGroupBudget) break;
    if (s->cfg.workspace_bytes && s->workspace_used + ws > s->cfg.workspace_bytes && s->workspace_used > 0) break;
    FPDB_TRY(ctx_alloc(c, ws, reinterpret_cast<void**>(&s->slot_partial[free_slot])));
    s->workspace_used += ws;
    R.workspace = ws;
    R.state = 1;
    used_groups += R.groups;
    s->slot_query[free_slot] = s->pending.front();
    s->pending.pop_front();
    admitted.push_back(free_slot);
  }
  return FPDB_OK;
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

// One shared scan for the admitted queries.  Holds the server's lock
// throughout, so submissions from other threads wait for the step.
static fpdb_status serve_step_impl(fpdb_serve* s, int* completed) {
  Context* c = s->ctx;
  *completed = 0;
  std::lock_guard<std::mutex> lock(s->mu);
  std::vector<int> admitted;
  FPDB_TRY(serve_admit(s, admitted));
  for (int slot : admitted) FPDB_TRY(serve_set_qmask(s, slot, s->queries[s->slot_query[slot]].q));
  ServeBatch B;
  std::memset(&B, 0, sizeof(B));
  B.ndims = s->ndims;
  for (int d = 0; d < s->ndims; ++d) B.dim[d] = s->dims[d];
  B.measure = s->measure;
  B.fact_rows = s->fact_rows;
  B.per_block = s->per_block;
  B.blocks = s->blocks;
  int nfin = 0, off = 0;
  for (int q = 0; q < kServeSlots; ++q) {
    if (s->slot_query[q] < 0) continue;
    const ServeQueryRec& R = s->queries[s->slot_query[q]];
    ServeSlot& S = B.slot[q];
    S.ngroup = R.q.num_group;
    S.gdim[0] = R.q.num_group >= 1 ? R.q.group_dim[0] : 0;
    S.gdim[1] = R.q.num_group >= 2 ? R.q.group_dim[1] : 0;
    S.dom1 = R.q.num_group >= 2 ? s->dims[S.gdim[1]].domain : 1;
    S.groups = R.groups;
    S.lds_off = off;
    off += R.groups;
    S.partial = s->slot_partial[q];
    S.out_g0 = s->res_g0[q];
    S.out_g1 = s->res_g1[q];
    S.out_cnt = s->res_cnt[q];
    S.out_sum = s->res_sum[q];
    S.out_n = s->res_n + q;
    B.active |= 1 << q;
    for (int f = 0; f < R.q.num_filters; ++f) B.users[R.q.filter_dim[f]] |= 1u << q;
    for (int i = 0; i < R.q.num_group; ++i) B.users[R.q.group_dim[i]] |= 1u << q;
    B.finalize_slot[nfin++] = q;
  }
  if (!nfin) return FPDB_OK;
  B.total_groups = off;
  {
    ProfScope ps(c, "serve_scan");
    hipLaunchKernelGGL(serve_scan_kernel, dim3(s->blocks), dim3(kServeThreads), (size_t)off * 16, c->stream, B);
  }
  FPDB_HIP(hipGetLastError());
  hipLaunchKernelGGL(serve_finalize_kernel, dim3(nfin), dim3(kServeThreads), 0, c->stream, B);
  FPDB_HIP(hipGetLastError());
  unsigned long long counts[kServeSlots];
  FPDB_HIP(hipMemcpyAsync(counts, s->res_n, sizeof(counts), hipMemcpyDeviceToHost, c->stream));
  FPDB_HIP(hipStreamSynchronize(c->stream));
  for (int q = 0; q < kServeSlots; ++q) {
    if (s->slot_query[q] < 0) continue;
    ServeQueryRec& R = s->queries[s->slot_query[q]];
    size_t n = (size_t)counts[q];
    R.g0.resize(n);
    R.g1.resize(n);
    R.cnt.resize(n);
    R.sum.resize(n);
    if (n) {
      FPDB_HIP(hipMemcpyAsync(R.g0.data(), s->res_g0[q], n * 4, hipMemcpyDeviceToHost, c->stream));
      FPDB_HIP(hipMemcpyAsync(R.g1.data(), s->res_g1[q], n * 4, hipMemcpyDeviceToHost, c->stream));
      FPDB_HIP(hipMemcpyAsync(R.cnt.data(), s->res_cnt[q], n * 8, hipMemcpyDeviceToHost, c->stream));
      FPDB_HIP(hipMemcpyAsync(R.sum.data(), s->res_sum[q], n * 8, hipMemcpyDeviceToHost, c->stream));
    }
  }
  FPDB_HIP(hipStreamSynchronize(c->stream));
  for (int q = 0; q < kServeSlots; ++q) {
    if (s->slot_query[q] < 0) continue;
    ServeQueryRec& R = s->queries[s->slot_query[q]];
    R.state = 2;
    ctx_free(c, s->slot_partial[q]);
    s->slot_partial[q] = nullptr;
    s->workspace_used -= R.workspace;
    s->slot_query[q] = -1;
    ++*completed;
  }
  ++s->steps;
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
  for (int q = 0; q < kServeSlots; ++q) {
    s->slot_query[q] = -1;
    s->slot_partial[q] = nullptr;
    s->res_g0[q] = s->res_g1[q] = nullptr;
    s->res_cnt[q] = s->res_sum[q] = nullptr;
  }
  std::memset(s->dims, 0, sizeof(s->dims));
  s->blocks = std::max(1, c->num_cus);
  s->per_block = std::max<int64_t>(1, ceil_div(std::max<int64_t>(s->fact_rows, 1), s->blocks));
  fpdb_status st = FPDB_OK;
  int* d_bad = nullptr;
  st = ctx_alloc(c, 8, reinterpret_cast<void**>(&d_bad));
  if (st == FPDB_OK) st = launch_fill(c, d_bad, 0, 8);
  for (int d = 0; d < s->ndims && st == FPDB_OK; ++d) {
    const fpdb_serve_dimension& D = schema->dims[d];
    ServeDim& X = s->dims[d];
    X.fk = static_cast<const int64_t*>(D.fact_fk.values);
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
  for (int q = 0; q < kServeSlots && st == FPDB_OK; ++q) {
    st = ctx_alloc(c, kServeMaxGroups * 4, reinterpret_cast<void**>(&s->res_g0[q]));
    if (st == FPDB_OK) st = ctx_alloc(c, kServeMaxGroups * 4, reinterpret_cast<void**>(&s->res_g1[q]));
    if (st == FPDB_OK) st = ctx_alloc(c, kServeMaxGroups * 8, reinterpret_cast<void**>(&s->res_cnt[q]));
    if (st == FPDB_OK) st = ctx_alloc(c, kServeMaxGroups * 8, reinterpret_cast<void**>(&s->res_sum[q]));
  }
  if (st == FPDB_OK) st = ctx_alloc(c, kServeSlots * 8, reinterpret_cast<void**>(&s->res_n));
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
// arrival time has passed, steps run while work is queued or in a scan, and
// each query's latency is its completion time minus its arrival time.
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
  std::vector<int64_t> index_of;   // query id -> trace index
  while (done < n) {
    int64_t t = now_ns();
    while (next < n && arrival_ns[next] <= t) {
      FPDB_TRY(fpdb_serve_submit(s, &queries[next], &query_ids[next]));
      if ((int64_t)index_of.size() <= query_ids[next]) index_of.resize(query_ids[next] + 1, -1);
      index_of[query_ids[next]] = next;
      ++next;
    }
    bool busy = false;
    {
      std::lock_guard<std::mutex> lock(s->mu);
      busy = !s->pending.empty();
    }
    if (!busy) {
      if (next < n) {
        int64_t wait = arrival_ns[next] - now_ns();
        if (wait > 200000) std::this_thread::sleep_for(std::chrono::nanoseconds(wait - 100000));
        continue;
      }
      return fail(FPDB_ERR_INTERNAL, "serve: trace stalled");
    }
    // the queries admitted by this step finish with it
    std::vector<int64_t> in_step;
    {
      std::lock_guard<std::mutex> lock(s->mu);
      for (size_t k = 0; k < s->pending.size() && (int)k < s->cfg.max_active; ++k) in_step.push_back(s->pending[k]);
    }
    int completed = 0;
    FPDB_TRY(serve_step_impl(s, &completed));
    int64_t t_done = now_ns();
    std::lock_guard<std::mutex> lock(s->mu);
    for (int64_t id : in_step) {
      if (s->queries[id].state != 2) continue;
      int64_t i = (id < (int64_t)index_of.size()) ? index_of[id] : -1;
      if (i >= 0) {
        latency_ns[i] = t_done - arrival_ns[i];
        ++done;
      }
    }
  }
  if (makespan_ns) *makespan_ns = now_ns();
  return FPDB_OK;
}

}  // extern "C"
