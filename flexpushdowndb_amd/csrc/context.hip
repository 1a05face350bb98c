// Context, device memory accounting, copies, timing and column descriptors.
#include "fpdb_internal.hpp"

#include <chrono>

#include <algorithm>

namespace fpdb {

static thread_local std::string g_last_error;

// Scratch of the no-allocation entry points (fpdb_filter_compact_noalloc):
// every GPU the library runs on holds its own copy until the process ends.
// Host code never uses this name as an address: the per-device address comes
// from hipGetSymbolAddress (noalloc_scratch_address).
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static std::unordered_map<int, uint8_t*> g_scratch_addr;

// Address of `device`'s copy of the scratch array; the device must be current.
static fpdb_status noalloc_scratch_address(int device, uint8_t** out) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  auto it = g_scratch_addr.find(device);
  if (it != g_scratch_addr.end()) {
    *out = it->second;
    return FPDB_OK;
  }
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr)
    return fail(FPDB_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
  g_scratch_addr[device] = static_cast<uint8_t*>(p);
  *out = static_cast<uint8_t*>(p);
  return FPDB_OK;
}

void set_error(const std::string& msg) { g_last_error = msg; }

fpdb_status fail(fpdb_status code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int type_width(int32_t t) {
  switch (t) {
    case FPDB_TYPE_INT8: case FPDB_TYPE_UINT8: return 1;
    case FPDB_TYPE_INT16: case FPDB_TYPE_UINT16: return 2;
    case FPDB_TYPE_INT32: case FPDB_TYPE_UINT32: case FPDB_TYPE_FLOAT32: case FPDB_TYPE_DATE32: return 4;
    case FPDB_TYPE_INT64: case FPDB_TYPE_UINT64: case FPDB_TYPE_FLOAT64: case FPDB_TYPE_DATE64: return 8;
    default: return 0;
  }
}

bool type_is_float(int32_t t) { return t == FPDB_TYPE_FLOAT32 || t == FPDB_TYPE_FLOAT64; }

bool type_is_signed_int(int32_t t) {
  return t == FPDB_TYPE_INT8 || t == FPDB_TYPE_INT16 || t == FPDB_TYPE_INT32 || t == FPDB_TYPE_INT64 ||
         t == FPDB_TYPE_DATE32 || t == FPDB_TYPE_DATE64;
}

fpdb_status make_colset(const fpdb_column* cols, int n, int64_t num_rows, ColSet* out,
                        const char* what) {
  FPDB_CHECK(n >= 0 && n <= kMaxCols, FPDB_ERR_INVALID,
             std::string(what) + ": too many columns (max " + std::to_string(kMaxCols) + ")");
  FPDB_CHECK(n == 0 || cols != nullptr, FPDB_ERR_INVALID, std::string(what) + ": columns is NULL");
  std::memset(out, 0, sizeof(*out));
  out->n = n;
  for (int i = 0; i < n; ++i) {
    int w = type_width(cols[i].type);
    FPDB_CHECK(w > 0, FPDB_ERR_INVALID,
               std::string(what) + ": column " + std::to_string(i) + " has unsupported type " +
                   std::to_string(cols[i].type));
    FPDB_CHECK(cols[i].length >= num_rows, FPDB_ERR_INVALID,
               std::string(what) + ": column " + std::to_string(i) + " has " +
                   std::to_string(cols[i].length) + " rows, " + std::to_string(num_rows) + " needed");
    FPDB_CHECK(num_rows == 0 || cols[i].values != nullptr, FPDB_ERR_INVALID,
               std::string(what) + ": column " + std::to_string(i) + " has no values buffer");
    FPDB_CHECK(cols[i].offset >= 0, FPDB_ERR_INVALID, std::string(what) + ": negative offset");
    out->values[i] = static_cast<const uint8_t*>(cols[i].values);
    out->validity[i] = cols[i].validity;
    out->offset[i] = cols[i].offset;
    out->width[i] = w;
    out->type[i] = cols[i].type;
  }
  return FPDB_OK;
}

fpdb_status make_outcolset(const fpdb_column* cols, int n, OutColSet* out, const char* what) {
  FPDB_CHECK(n >= 0 && n <= kMaxCols, FPDB_ERR_INVALID, std::string(what) + ": too many columns");
  FPDB_CHECK(n == 0 || cols != nullptr, FPDB_ERR_INVALID, std::string(what) + ": columns is NULL");
  std::memset(out, 0, sizeof(*out));
  out->n = n;
  out->capacity = INT64_MAX;
  for (int i = 0; i < n; ++i) {
    int w = type_width(cols[i].type);
    FPDB_CHECK(w > 0, FPDB_ERR_INVALID,
               std::string(what) + ": output column " + std::to_string(i) + " has unsupported type");
    FPDB_CHECK(cols[i].length >= 0, FPDB_ERR_INVALID,
               std::string(what) + ": output column " + std::to_string(i) + " has a negative length");
    out->capacity = std::min<int64_t>(out->capacity, cols[i].length);
    out->values[i] = static_cast<uint8_t*>(cols[i].values);
    out->validity[i] = cols[i].validity;
    out->offset[i] = cols[i].offset;
    out->width[i] = w;
  }
  return FPDB_OK;
}

// Allocation sizes are rounded (powers of two up to 1 MiB, then 1 MiB
// multiples) so freed blocks can be reused by later calls of similar size.
static uint64_t round_alloc(uint64_t bytes) {
  if (bytes <= 256) return 256;
  if (bytes <= (1ULL << 20)) {
    uint64_t r = 256;
    while (r < bytes) r <<= 1;
    return r;
  }
  return (bytes + (1ULL << 20) - 1) & ~((1ULL << 20) - 1);
}

static void release_cache_locked(Context* c) {
  for (auto& kv : c->cache) (void)hipFree(kv.second);
  c->cache.clear();
  c->cached = 0;
}

fpdb_status ctx_alloc(Context* c, uint64_t bytes, void** out) {
  *out = nullptr;
  uint64_t r = round_alloc(bytes == 0 ? 1 : bytes);
  std::lock_guard<std::mutex> g(c->mu);
  if (c->budget && c->used + r > c->budget)
    return fail(FPDB_ERR_BUDGET, "HBM budget exceeded: " + std::to_string(c->used) + " + " +
                                     std::to_string(r) + " > " + std::to_string(c->budget));
  auto it = c->cache.find(r);
  if (it != c->cache.end()) {
    *out = it->second;
    c->cache.erase(it);
    c->cached -= r;
  } else {
    hipError_t e = hipMalloc(out, r);
    if (e != hipSuccess && c->cached) {
      (void)hipGetLastError();
      release_cache_locked(c);
      e = hipMalloc(out, r);
    }
    if (e != hipSuccess) {
      *out = nullptr;
      return fail(FPDB_ERR_HIP, std::string("hipMalloc(") + std::to_string(r) + ") failed: " +
                                    hipGetErrorString(e));
    }
  }
  c->used += r;
  c->allocs[*out] = r;
  return FPDB_OK;
}

// Blocks return to the context's cache; work already enqueued on the stream
// that still reads them is ordered before any later reuse on the same stream.
fpdb_status ctx_free(Context* c, void* p) {
  if (!p) return FPDB_OK;
  std::lock_guard<std::mutex> g(c->mu);
  auto it = c->allocs.find(p);
  FPDB_CHECK(it != c->allocs.end(), FPDB_ERR_INVALID, "fpdb_device_free: pointer not owned by context");
  c->used -= it->second;
  c->cache.emplace(it->second, p);
  c->cached += it->second;
  c->allocs.erase(it);
  return FPDB_OK;
}

// One wave: the values, then (after a system-scope fence that waits for the
// wave's stores) the sequence number.
__global__ void mail_kernel(const uint64_t* __restrict__ src, int nwords, uint64_t* mail, uint64_t seq) {
  const int lane = threadIdx.x;
  if (lane < nwords) __hip_atomic_store(mail + 1 + lane, src[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __threadfence_system();
  if (lane == 0) __hip_atomic_store(mail, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

fpdb_status mail_read(Context* c, const void* d_src, int nwords, uint64_t* out) {
  FPDB_CHECK(nwords >= 0 && nwords <= 7, FPDB_ERR_INTERNAL, "mail_read: too many words");
  const uint64_t seq = ++c->mail_seq;
  hipLaunchKernelGGL(mail_kernel, dim3(1), dim3(64), 0, c->stream, static_cast<const uint64_t*>(d_src), nwords,
                     c->d_mail, seq);
  FPDB_HIP(hipGetLastError());
  const auto t0 = std::chrono::steady_clock::now();
  for (uint64_t i = 0;; ++i) {
    if (__atomic_load_n(c->h_mail, __ATOMIC_ACQUIRE) == seq) break;
    if ((i & 4095) == 4095 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(50)) {
      FPDB_HIP(hipStreamSynchronize(c->stream));
      FPDB_CHECK(__atomic_load_n(c->h_mail, __ATOMIC_ACQUIRE) == seq, FPDB_ERR_INTERNAL,
                 "read-back did not arrive after the stream completed");
      break;
    }
  }
  for (int i = 0; i < nwords; ++i) out[i] = __atomic_load_n(c->h_mail + 1 + i, __ATOMIC_RELAXED);
  return FPDB_OK;
}

fpdb_status read_count(Context* c, int64_t* count) {
  uint64_t v[2];
  FPDB_TRY(mail_read(c, c->d_counter, 2, v));
  if ((int)(uint32_t)v[1] != 0) {
    FPDB_TRY(launch_fill(c, c->d_counter + 1, 0, 8));
    return fail(FPDB_ERR_INTERNAL, "a kernel reported a failure (a bounded wait gave up or a table overflowed)");
  }
  if (count) *count = (int64_t)v[0];
  return FPDB_OK;
}

hipEvent_t prof_event(Context* c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

__global__ void fill_kernel(uint8_t* __restrict__ d, uint8_t v, uint64_t bytes) {
  uint64_t v8 = 0x0101010101010101ULL * v;
  uint64_t head = (8 - (reinterpret_cast<uintptr_t>(d) & 7)) & 7;
  if (head > bytes) head = bytes;
  uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  if (tid < head) d[tid] = v;
  uint64_t words = (bytes - head) / 8;
  uint64_t* w = reinterpret_cast<uint64_t*>(d + head);
  for (uint64_t i = tid; i < words; i += stride) w[i] = v8;
  uint64_t tail_start = head + words * 8;
  if (tid < bytes - tail_start) d[tail_start + tid] = v;
}

fpdb_status launch_fill(Context* c, void* d, uint8_t v, uint64_t bytes) {
  if (bytes == 0) return FPDB_OK;
  FPDB_CHECK(d != nullptr, FPDB_ERR_INVALID, "fill: NULL destination");
  uint64_t words = bytes / 8 + 1;
  int blocks = (int)std::min<uint64_t>((words + 255) / 256, 4096);
  hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, c->stream,
                     static_cast<uint8_t*>(d), v, bytes);
  FPDB_HIP(hipGetLastError());
  return FPDB_OK;
}

__global__ __launch_bounds__(1024) void sum_partials_kernel(const unsigned long long* __restrict__ partials,
                                                             int n, int64_t* dst) {
  __shared__ unsigned long long s_w[16];
  unsigned long long v = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) v += partials[i];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += s_w[w];
    *dst = (int64_t)t;
  }
}

fpdb_status launch_sum_partials(Context* c, int blocks, int64_t* dst) {
  FPDB_CHECK(blocks >= 1 && blocks <= kMaxPartials, FPDB_ERR_INTERNAL, "too many workgroups for the partial counts");
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(1024), 0, c->stream, c->d_partials, blocks, dst);
  FPDB_HIP(hipGetLastError());
  return FPDB_OK;
}

}  // namespace fpdb

using namespace fpdb;

extern "C" {

const char* fpdb_last_error(void) { return g_last_error.c_str(); }

fpdb_status fpdb_context_create(int device, uint64_t hbm_budget_bytes, void* stream,
                                fpdb_context** out) {
  FPDB_CHECK(out != nullptr, FPDB_ERR_INVALID, "fpdb_context_create: out is NULL");
  *out = nullptr;
  int ndev = 0;
  FPDB_HIP(hipGetDeviceCount(&ndev));
  FPDB_CHECK(device >= 0 && device < ndev, FPDB_ERR_INVALID,
             "fpdb_context_create: device " + std::to_string(device) + " not present (" +
                 std::to_string(ndev) + " devices)");
  FPDB_HIP(hipSetDevice(device));
  Context* c = new Context();
  c->device = device;
  c->budget = hbm_budget_bytes;
  if (stream) {
    c->stream = static_cast<hipStream_t>(stream);
  } else {
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      return fail(FPDB_ERR_HIP, std::string("hipStreamCreate failed: ") + hipGetErrorString(e));
    }
    c->own_stream = true;
  }
  hipDeviceProp_t prop;
  FPDB_HIP(hipGetDeviceProperties(&prop, device));
  {
    fpdb_status s = noalloc_scratch_address(device, &c->noalloc_scratch);
    if (s != FPDB_OK) {
      if (c->own_stream) (void)hipStreamDestroy(c->stream);
      delete c;
      return s;
    }
  }
  c->num_cus = prop.multiProcessorCount;
  c->hbm_bytes = prop.totalGlobalMem;
  c->max_lds_per_block = prop.sharedMemPerBlock;
  c->num_xcds = 8;
  FPDB_HIP(hipEventCreate(&c->ev_start));
  FPDB_HIP(hipEventCreate(&c->ev_stop));
  fpdb_status s = ctx_alloc(c, 64, reinterpret_cast<void**>(&c->d_counter));
  if (s == FPDB_OK) s = ctx_alloc(c, (uint64_t)kMaxPartials * 8, reinterpret_cast<void**>(&c->d_partials));
  if (s == FPDB_OK) s = launch_fill(c, c->d_counter, 0, 64);
  if (s != FPDB_OK) { delete c; return s; }
  FPDB_HIP(hipHostMalloc(reinterpret_cast<void**>(&c->h_counter), 64, hipHostMallocDefault));
  FPDB_HIP(hipHostMalloc(reinterpret_cast<void**>(&c->h_mail), 64, hipHostMallocCoherent | hipHostMallocMapped));
  std::memset(c->h_mail, 0, 64);
  FPDB_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&c->d_mail), c->h_mail, 0));
  *out = reinterpret_cast<fpdb_context*>(c);
  return FPDB_OK;
}

fpdb_status fpdb_context_destroy(fpdb_context* ctx) {
  if (!ctx) return FPDB_OK;
  Context* c = as_ctx(ctx);
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto& kv : c->allocs) (void)hipFree(kv.first);
  c->allocs.clear();
  release_cache_locked(c);
  for (auto& r : c->prof) {
    (void)hipEventDestroy(r.start);
    (void)hipEventDestroy(r.stop);
  }
  for (auto e : c->event_pool) (void)hipEventDestroy(e);
  if (c->h_counter) (void)hipHostFree(c->h_counter);
  if (c->h_mail) (void)hipHostFree(c->h_mail);
  if (c->ev_start) (void)hipEventDestroy(c->ev_start);
  if (c->ev_stop) (void)hipEventDestroy(c->ev_stop);
  if (c->own_stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return FPDB_OK;
}

void* fpdb_context_stream(fpdb_context* ctx) { return ctx ? as_ctx(ctx)->stream : nullptr; }

fpdb_status fpdb_context_synchronize(fpdb_context* ctx) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  Context* c = as_ctx(ctx);
  FPDB_HIP(hipStreamSynchronize(c->stream));
  int flag = 0;
  FPDB_HIP(hipMemcpy(&flag, ctx_err_flag(c), sizeof(int), hipMemcpyDeviceToHost));
  if (flag) {
    FPDB_TRY(launch_fill(c, ctx_err_flag(c), 0, 8));
    FPDB_HIP(hipStreamSynchronize(c->stream));
    return fail(FPDB_ERR_INTERNAL, "a kernel reported a failure (a bounded wait gave up or a table overflowed)");
  }
  return FPDB_OK;
}

fpdb_status fpdb_context_memory(fpdb_context* ctx, uint64_t* used, uint64_t* budget) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  Context* c = as_ctx(ctx);
  std::lock_guard<std::mutex> g(c->mu);
  if (used) *used = c->used;
  if (budget) *budget = c->budget;
  return FPDB_OK;
}

fpdb_status fpdb_device_info(fpdb_context* ctx, int* num_cus, int* num_xcds, uint64_t* hbm_bytes) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  Context* c = as_ctx(ctx);
  if (num_cus) *num_cus = c->num_cus;
  if (num_xcds) *num_xcds = c->num_xcds;
  if (hbm_bytes) *hbm_bytes = c->hbm_bytes;
  return FPDB_OK;
}

fpdb_status fpdb_device_alloc(fpdb_context* ctx, uint64_t bytes, void** d_out) {
  FPDB_CHECK(ctx && d_out, FPDB_ERR_INVALID, "fpdb_device_alloc: NULL argument");
  return ctx_alloc(as_ctx(ctx), bytes, d_out);
}

fpdb_status fpdb_device_free(fpdb_context* ctx, void* d_ptr) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  return ctx_free(as_ctx(ctx), d_ptr);
}

fpdb_status fpdb_copy_to_device(fpdb_context* ctx, void* d_dst, const void* src, uint64_t bytes) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  if (bytes == 0) return FPDB_OK;
  FPDB_HIP(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, as_ctx(ctx)->stream));
  FPDB_HIP(hipStreamSynchronize(as_ctx(ctx)->stream));
  return FPDB_OK;
}

fpdb_status fpdb_copy_to_host(fpdb_context* ctx, void* dst, const void* d_src, uint64_t bytes) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  if (bytes == 0) return FPDB_OK;
  FPDB_HIP(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, as_ctx(ctx)->stream));
  FPDB_HIP(hipStreamSynchronize(as_ctx(ctx)->stream));
  return FPDB_OK;
}

fpdb_status fpdb_copy_device(fpdb_context* ctx, void* d_dst, const void* d_src, uint64_t bytes) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  if (bytes == 0) return FPDB_OK;
  FPDB_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, as_ctx(ctx)->stream));
  return FPDB_OK;
}

fpdb_status fpdb_fill(fpdb_context* ctx, void* d_dst, uint8_t value, uint64_t bytes) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  return launch_fill(as_ctx(ctx), d_dst, value, bytes);
}

fpdb_status fpdb_context_set_profiling(fpdb_context* ctx, int enabled) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  as_ctx(ctx)->profiling = enabled != 0;
  return FPDB_OK;
}

fpdb_status fpdb_profile_kernel(fpdb_context* ctx, const char* name, double* total_ms, int64_t* launches) {
  FPDB_CHECK(ctx && name && total_ms && launches, FPDB_ERR_INVALID, "NULL argument");
  Context* c = as_ctx(ctx);
  FPDB_HIP(hipStreamSynchronize(c->stream));
  double t = 0;
  int64_t n = 0;
  for (auto& r : c->prof) {
    if (std::strcmp(r.name, name) != 0) continue;
    float ms = 0;
    FPDB_HIP(hipEventElapsedTime(&ms, r.start, r.stop));
    t += ms;
    ++n;
  }
  *total_ms = t;
  *launches = n;
  return FPDB_OK;
}

fpdb_status fpdb_profile_reset(fpdb_context* ctx) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  Context* c = as_ctx(ctx);
  FPDB_HIP(hipStreamSynchronize(c->stream));
  for (auto& r : c->prof) {
    c->event_pool.push_back(r.start);
    c->event_pool.push_back(r.stop);
  }
  c->prof.clear();
  return FPDB_OK;
}

fpdb_status fpdb_timer_start(fpdb_context* ctx) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  FPDB_HIP(hipEventRecord(as_ctx(ctx)->ev_start, as_ctx(ctx)->stream));
  return FPDB_OK;
}

fpdb_status fpdb_timer_stop(fpdb_context* ctx, float* ms) {
  FPDB_CHECK(ctx && ms, FPDB_ERR_INVALID, "NULL argument");
  Context* c = as_ctx(ctx);
  FPDB_HIP(hipEventRecord(c->ev_stop, c->stream));
  FPDB_HIP(hipEventSynchronize(c->ev_stop));
  FPDB_HIP(hipEventElapsedTime(ms, c->ev_start, c->ev_stop));
  return FPDB_OK;
}

}  // extern "C"
