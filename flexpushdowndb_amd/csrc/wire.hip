// Wire forms of the filters and selection bitmaps that the reference ships to
// its store (Flight DoPut of Arrow record batches):
//   * a selection bitmap as int64 words, bit r % 64 of word r / 64 = row r
//     (fpdb::util::setBit / getBit, fpdb-util/src/Util.cpp:115-125), filled
//     batch after batch at a running row offset
//     (FilterPOp::bufferBitMap, fpdb-executor/src/physical/filter/FilterPOp.cpp:264-279)
//     and read back per batch (FilterPOp::makeSelectionVector, :281-297);
//   * a blocked bloom filter as the 136 mask bytes and the block bytes
//     (ArrowBloomFilter::makeBitmapRecordBatches / saveBitmapRecordBatches,
//     fpdb-executor/src/physical/bloomfilter/ArrowBloomFilter.cpp:77-116).
// The library's selection bitmaps are LSB-first bytes from row 0, so the
// conversions are shifted bit copies: one thread per destination word.
#include "bloom_masks.hpp"
#include "fpdb_internal.hpp"

namespace fpdb {

// Bits [s, s + 64) of an LSB-first byte bitmap of nbytes bytes (bits outside
// the bytes read as 0; s may be negative).
__device__ __forceinline__ uint64_t bits64_from_bytes(const uint8_t* b, int64_t nbytes, int64_t s) {
  const int64_t byte0 = s >= 0 ? s >> 3 : -((-s + 7) >> 3);
  const int sh = (int)(s - byte0 * 8);   // 0..7
  uint64_t lo = 0;
  uint32_t hi = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int64_t i = byte0 + j;
    if (i >= 0 && i < nbytes) lo |= (uint64_t)b[i] << (8 * j);
  }
  const int64_t i8 = byte0 + 8;
  if (i8 >= 0 && i8 < nbytes) hi = b[i8];
  return sh ? (lo >> sh) | ((uint64_t)hi << (64 - sh)) : lo;
}

// words[k] |= the bitmap's bits for rows [64k - row_offset, ...) that fall in [0, n).
__global__ void bitmap_to_words_kernel(const uint8_t* __restrict__ bm, int64_t n, int64_t row_offset,
                                       int64_t k0, int64_t k1, uint64_t* __restrict__ words) {
  const int64_t nbytes = (n + 7) / 8;
  for (int64_t k = k0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < k1;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = k * 64 - row_offset;           // bitmap bit of the word's bit 0
    uint64_t v = bits64_from_bytes(bm, nbytes, s);
    // keep only bits for rows [0, n) of this bitmap
    const int64_t lo = s < 0 ? -s : 0;                 // first word bit holding row 0
    const int64_t hi = n - s;                          // word bits >= hi hold rows >= n
    if (lo > 0) v &= lo >= 64 ? 0 : ~0ULL << lo;
    if (hi < 64) v &= hi <= 0 ? 0 : (~0ULL >> (64 - hi));
    words[k] |= v;
  }
}

// Selection bitmap of rows [row_offset, row_offset + n): (n + 7) / 8 bytes.
__global__ void words_to_bitmap_kernel(const uint64_t* __restrict__ words, int64_t num_words, int64_t row_offset,
                                       int64_t n, uint8_t* __restrict__ bm) {
  const int64_t chunks = (n + 63) / 64;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < chunks; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = row_offset + q * 64;
    const int64_t w = s >> 6;
    const int sh = (int)(s & 63);
    uint64_t v = w < num_words ? words[w] >> sh : 0;
    if (sh && w + 1 < num_words) v |= words[w + 1] << (64 - sh);
    const int64_t rem = n - q * 64;
    if (rem < 64) v &= (~0ULL >> (64 - rem));
    const int nb = rem >= 64 ? 8 : (int)((rem + 7) / 8);
    for (int j = 0; j < nb; ++j) bm[q * 8 + j] = (uint8_t)(v >> (8 * j));
  }
}

}  // namespace fpdb

using namespace fpdb;

extern "C" fpdb_status fpdb_bitmap_to_words(fpdb_context* ctx, const uint8_t* d_bitmap, int64_t n, int64_t row_offset,
                                            int64_t* d_words) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  FPDB_CHECK(n >= 0 && row_offset >= 0, FPDB_ERR_INVALID, "fpdb_bitmap_to_words: negative rows or offset");
  if (n == 0) return FPDB_OK;
  FPDB_CHECK(d_bitmap && d_words, FPDB_ERR_INVALID, "fpdb_bitmap_to_words: NULL buffer");
  Context* c = as_ctx(ctx);
  const int64_t k0 = row_offset / 64, k1 = (row_offset + n + 63) / 64;
  const int64_t words = k1 - k0;
  const int grid = (int)std::min<int64_t>(std::max<int64_t>(1, (words + 255) / 256), (int64_t)c->num_cus * 8);
  hipLaunchKernelGGL(bitmap_to_words_kernel, dim3(grid), dim3(256), 0, c->stream, d_bitmap, n, row_offset, k0, k1,
                     reinterpret_cast<uint64_t*>(d_words));
  FPDB_HIP(hipGetLastError());
  return FPDB_OK;
}

extern "C" fpdb_status fpdb_bitmap_from_words(fpdb_context* ctx, const int64_t* d_words, int64_t num_words,
                                              int64_t row_offset, int64_t n, uint8_t* d_bitmap) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  FPDB_CHECK(n >= 0 && row_offset >= 0 && num_words >= 0, FPDB_ERR_INVALID,
             "fpdb_bitmap_from_words: negative rows, offset or word count");
  FPDB_CHECK(row_offset + n <= num_words * 64, FPDB_ERR_INVALID,
             "fpdb_bitmap_from_words: rows past the end of the words");
  if (n == 0) return FPDB_OK;
  FPDB_CHECK(d_bitmap && d_words, FPDB_ERR_INVALID, "fpdb_bitmap_from_words: NULL buffer");
  Context* c = as_ctx(ctx);
  const int64_t chunks = (n + 63) / 64;
  const int grid = (int)std::min<int64_t>(std::max<int64_t>(1, (chunks + 255) / 256), (int64_t)c->num_cus * 8);
  hipLaunchKernelGGL(words_to_bitmap_kernel, dim3(grid), dim3(256), 0, c->stream,
                     reinterpret_cast<const uint64_t*>(d_words), num_words, row_offset, n, d_bitmap);
  FPDB_HIP(hipGetLastError());
  return FPDB_OK;
}

extern "C" fpdb_status fpdb_bloom_blocked_export(fpdb_context* ctx, const uint64_t* d_blocks, int log_num_blocks,
                                                 uint8_t* masks_out, uint8_t* blocks_out) {
  FPDB_CHECK(ctx, FPDB_ERR_INVALID, "NULL context");
  FPDB_CHECK(log_num_blocks >= 0 && log_num_blocks <= 40, FPDB_ERR_INVALID, "bad log_num_blocks");
  if (masks_out) FPDB_TRY(fpdb_bloom_masks(masks_out));
  if (blocks_out) {
    FPDB_CHECK(d_blocks, FPDB_ERR_INVALID, "fpdb_bloom_blocked_export: NULL blocks");
    FPDB_TRY(fpdb_copy_to_host(ctx, blocks_out, d_blocks, (uint64_t)8 << log_num_blocks));
  }
  return FPDB_OK;
}

extern "C" fpdb_status fpdb_bloom_blocked_import(fpdb_context* ctx, const uint8_t* masks, const uint8_t* blocks,
                                                 uint64_t block_bytes, int* log_num_blocks, uint64_t* d_blocks) {
  FPDB_CHECK(ctx && masks && blocks && log_num_blocks, FPDB_ERR_INVALID, "fpdb_bloom_blocked_import: NULL argument");
  uint8_t mine[136];
  FPDB_TRY(fpdb_bloom_masks(mine));
  FPDB_CHECK(std::memcmp(mine, masks, sizeof(mine)) == 0, FPDB_ERR_INVALID,
             "fpdb_bloom_blocked_import: the filter's masks differ from BloomFilterMasks()");
  FPDB_CHECK(block_bytes >= 8 && block_bytes % 8 == 0 && ((block_bytes / 8) & (block_bytes / 8 - 1)) == 0,
             FPDB_ERR_INVALID, "fpdb_bloom_blocked_import: block bytes must be 8 x a power of two");
  int lnb = 0;
  while (((uint64_t)8 << lnb) < block_bytes) ++lnb;
  *log_num_blocks = lnb;
  if (d_blocks) FPDB_TRY(fpdb_copy_to_device(ctx, d_blocks, blocks, block_bytes));
  return FPDB_OK;
}
