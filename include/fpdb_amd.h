/*
 * fpdb_amd.h -- C ABI of the MI355X (gfx950) relational hot path of
 * FlexPushdownDB: blocked and legacy bloom filters (predicate transfer),
 * hash-join build/probe, hash group-by, filter compaction, hash shuffle and
 * a device-resident segment cache.
 *
 * Conventions
 *   - Every entry point returns an fpdb_status; FPDB_OK is 0.  On failure
 *     fpdb_last_error() returns a message for the calling thread.  Nothing
 *     throws across this boundary.
 *   - Pointers named d_* are device (HBM) pointers; everything else is host.
 *   - Inputs are borrowed for the duration of the call and never mutated.
 *   - Work is enqueued on the context's stream; results are ready after
 *     fpdb_context_synchronize() unless a function says it synchronizes.
 *   - Columns use the Arrow layout: a values buffer, an optional validity
 *     bitmap (LSB-first, 1 = valid) and an element offset.
 *   - Calls on one context are serialised by the caller (one operator instance
 *     per context, like one actor per physical operator); different contexts
 *     may be used from different threads concurrently.
 *
 * Reference paths below are relative to the FlexPushdownDB source tree.
 */
#ifndef FPDB_AMD_H
#define FPDB_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t fpdb_status;
enum {
  FPDB_OK = 0,
  FPDB_ERR_INVALID = 1,     /* bad argument / unsupported type or layout   */
  FPDB_ERR_HIP = 2,         /* HIP runtime error                           */
  FPDB_ERR_BUDGET = 3,      /* allocation would exceed the HBM budget      */
  FPDB_ERR_CAPACITY = 4,    /* an output or table capacity was exceeded    */
  FPDB_ERR_INTERNAL = 5     /* kernel-side consistency check failed        */
};

/* Arrow type ids supported on the hot path. */
typedef enum {
  FPDB_TYPE_INT8 = 1, FPDB_TYPE_INT16 = 2, FPDB_TYPE_INT32 = 3, FPDB_TYPE_INT64 = 4,
  FPDB_TYPE_UINT8 = 5, FPDB_TYPE_UINT16 = 6, FPDB_TYPE_UINT32 = 7, FPDB_TYPE_UINT64 = 8,
  FPDB_TYPE_FLOAT32 = 9, FPDB_TYPE_FLOAT64 = 10, FPDB_TYPE_DATE32 = 11, FPDB_TYPE_DATE64 = 12,
  FPDB_TYPE_UTF8 = 13
} fpdb_type;

/* One Arrow-layout column on the device.  For outputs, values/validity are
 * caller-allocated buffers the library writes into (validity may be NULL when
 * the output cannot hold nulls).
 * A utf8 column (Arrow "utf8": int32 offsets + character bytes) keeps its
 * offsets in `values` (offset + length + 1 entries; row r spans bytes
 * [values[offset + r], values[offset + r + 1]) of `data`) and its bytes in
 * `data`; `data_bytes` is the size of `data` (for outputs: the room in it).
 * Entry points that take utf8 columns say so; the others reject them. */
typedef struct fpdb_column {
  int32_t type;             /* fpdb_type                                   */
  int32_t reserved;
  int64_t length;           /* rows                                        */
  int64_t offset;           /* element offset into values and validity     */
  void* values;             /* device pointer (utf8: int32 offsets)        */
  uint8_t* validity;        /* device pointer or NULL (all valid)          */
  void* data;               /* utf8: character bytes (device); else NULL   */
  int64_t data_bytes;       /* utf8: bytes in (room of) data; else 0       */
} fpdb_column;

typedef struct fpdb_context fpdb_context;

/* ---- context, memory, timing ------------------------------------------- */

/* Creates a context on `device` with an HBM budget (0 = unlimited) for the
 * buffers this library allocates.  If `stream` is NULL the context creates its
 * own default-priority stream.  Replaces the per-operator resources the
 * reference binds in POpContext (fpdb-executor/src/physical/POpContext.cpp:25-58). */
fpdb_status fpdb_context_create(int device, uint64_t hbm_budget_bytes, void* stream,
                                fpdb_context** out);
/* Destroy every object created on a context (tables, filters, group tables,
 * caches, star joins, servers) before the context itself: their destroy calls
 * return memory to it. */
fpdb_status fpdb_context_destroy(fpdb_context* ctx);
void* fpdb_context_stream(fpdb_context* ctx);
fpdb_status fpdb_context_synchronize(fpdb_context* ctx);
/* Bytes currently allocated through the context, and its budget. */
fpdb_status fpdb_context_memory(fpdb_context* ctx, uint64_t* used, uint64_t* budget);
/* Number of compute units and XCDs of the context's device. */
fpdb_status fpdb_device_info(fpdb_context* ctx, int* num_cus, int* num_xcds, uint64_t* hbm_bytes);
const char* fpdb_last_error(void);

fpdb_status fpdb_device_alloc(fpdb_context* ctx, uint64_t bytes, void** d_out);
fpdb_status fpdb_device_free(fpdb_context* ctx, void* d_ptr);
/* Stream-ordered copies; fpdb_copy_to_host synchronizes before returning. */
fpdb_status fpdb_copy_to_device(fpdb_context* ctx, void* d_dst, const void* src, uint64_t bytes);
fpdb_status fpdb_copy_to_host(fpdb_context* ctx, void* dst, const void* d_src, uint64_t bytes);
fpdb_status fpdb_copy_device(fpdb_context* ctx, void* d_dst, const void* d_src, uint64_t bytes);
/* Fills with a byte value using a kernel (safe inside captured graphs). */
fpdb_status fpdb_fill(fpdb_context* ctx, void* d_dst, uint8_t value, uint64_t bytes);

/* HIP events on the context stream: elapsed milliseconds between the last
 * fpdb_timer_start and fpdb_timer_stop (fpdb_timer_stop synchronizes). */
fpdb_status fpdb_timer_start(fpdb_context* ctx);
fpdb_status fpdb_timer_stop(fpdb_context* ctx, float* ms);

/* Kernel timing (the reference's per-operator processing time,
 * fpdb-executor/src/physical/POpActor.cpp:40-97, at kernel granularity): while
 * enabled, every named kernel launch is bracketed by HIP events on the context
 * stream.  fpdb_profile_kernel synchronizes and returns the summed duration
 * and launch count of one kernel name ("hj_build", "hj_probe",
 * "bloom_blocked_insert", "bloom_blocked_find", "bloom_blocked_filter",
 * "bloom_legacy_build", "bloom_legacy_find", "filter_compact",
 * "groupby_consume"). */
fpdb_status fpdb_context_set_profiling(fpdb_context* ctx, int enabled);
fpdb_status fpdb_profile_kernel(fpdb_context* ctx, const char* name, double* total_ms,
                                int64_t* launches);
fpdb_status fpdb_profile_reset(fpdb_context* ctx);

/* ---- key hashing -------------------------------------------------------- */

/* uint32 hash per row of fixed-width key columns, as the reference computes it
 * with Arrow 6 Hashing::HashMultiColumn in RecordBatchHasher::hash
 * (fpdb-tuple/src/RecordBatchHasher.cpp:69-104). */
fpdb_status fpdb_hash_keys(fpdb_context* ctx, const fpdb_column* keys, int num_keys,
                           int64_t num_rows, uint32_t* d_hashes);

/* ---- blocked bloom filter (default, Arrow BlockedBloomFilter) ----------- */

/* The 136-byte mask bit vector (BloomFilterMasks(),
 * fpdb-tuple/src/arrow/exec/BloomFilter.cpp:16-79). */
fpdb_status fpdb_bloom_masks(uint8_t out_masks[136]);

/* log2(number of 64-bit blocks) for a filter sized for `num_rows_to_insert`
 * keys (BlockedBloomFilter::CreateEmpty, BloomFilter.cpp:87-107). */
int fpdb_bloom_blocked_log_num_blocks(int64_t num_rows_to_insert);

/* Zeroes the (1 << log_num_blocks) blocks and inserts every key row
 * (ArrowBloomFilter::build, fpdb-executor/src/physical/bloomfilter/ArrowBloomFilter.cpp:29-75;
 * BlockedBloomFilter::Insert, BloomFilter.cpp:109-136).  If d_select is not
 * NULL only rows whose bit is set are inserted. */
fpdb_status fpdb_bloom_blocked_build(fpdb_context* ctx, const fpdb_column* keys, int num_keys,
                                     int64_t num_rows, const uint8_t* d_select,
                                     int log_num_blocks, uint64_t* d_blocks);

/* Inserts into an existing filter without clearing it (a filter of several
 * batches, or the OR-merge of partitioned builds). */
fpdb_status fpdb_bloom_blocked_insert(fpdb_context* ctx, const fpdb_column* keys, int num_keys,
                                      int64_t num_rows, const uint8_t* d_select,
                                      int log_num_blocks, uint64_t* d_blocks);

/* Packed selection bitmap (bit i = Find(hash(row i)), LSB-first, (n+7)/8
 * bytes written) and, if d_num_selected is not NULL, the number of selected
 * rows (stored as int64 on the device)
 * (BlockedBloomFilter::Find, BloomFilter.cpp:138-191). */
fpdb_status fpdb_bloom_blocked_find(fpdb_context* ctx, const uint64_t* d_blocks, int log_num_blocks,
                                    const fpdb_column* keys, int num_keys, int64_t num_rows,
                                    uint8_t* d_bitmap, int64_t* d_num_selected);

/* ---- legacy bloom filter (universal hashing, target FPR) ---------------- */

/* One Carter-Wegman hash h(x) = ((a*x + b) mod p) mod m
 * (UniversalHashFunction, fpdb-executor/src/physical/bloomfilter/UniversalHashFunction.cpp:51-57). */
typedef struct fpdb_universal_hash { int64_t a, b, m, p; } fpdb_universal_hash;

/* k = ceil(log2(1/p)), m = ceil(n |ln p| / ln^2 2)
 * (BloomFilter::k_from_p / m_from_np, BloomFilter.cpp:249-257). */
fpdb_status fpdb_bloom_legacy_params(int64_t capacity, double false_positive_rate,
                                     int64_t* num_hash_functions, int64_t* num_bits);

/* Universal-hash parameters drawn like UniversalHashFunction(m)
 * (UniversalHashFunction.cpp:13-35) from an explicit seed instead of
 * std::random_device, so the filter is reproducible. */
fpdb_status fpdb_bloom_legacy_make_hash_functions(uint32_t seed, int64_t num_bits, int k,
                                                  fpdb_universal_hash* out);

/* Zeroes num_bits/64+1 words and adds every key row
 * (BloomFilterCreateKernel::buildBloomFilter / addRecordBatchToBloomFilter,
 * BloomFilterCreateKernel.cpp:24-99; BloomFilter::add, BloomFilter.cpp:62-73). */
fpdb_status fpdb_bloom_legacy_build(fpdb_context* ctx, const fpdb_column* keys, int num_keys,
                                    int64_t num_rows, const uint8_t* d_select,
                                    const fpdb_universal_hash* hash_functions, int k,
                                    int64_t num_bits, int64_t* d_words);

/* Selection bitmap of BloomFilter::contains (BloomFilter.cpp:75-91) per row,
 * i.e. the rows the legacy filterRecordBatch keeps (BloomFilterUseKernel.cpp:140-178).
 * capacity == 0 selects nothing. */
fpdb_status fpdb_bloom_legacy_find(fpdb_context* ctx, const fpdb_universal_hash* hash_functions,
                                   int k, int64_t capacity, int64_t num_bits, const int64_t* d_words,
                                   const fpdb_column* keys, int num_keys, int64_t num_rows,
                                   uint8_t* d_bitmap, int64_t* d_num_selected);

/* ---- filter compaction --------------------------------------------------- */

/* Comparison of one column with a constant into a selection bitmap (nulls are
 * not selected), AND-ed with d_and_bitmap when given, or OR-ed with it when op
 * carries FPDB_CMP_OR (conjunctions and disjunctions of such comparisons chain
 * through the bitmap; d_and_bitmap may be d_bitmap itself).  Integer columns
 * compare with int_constant, float columns with float_constant.  Stands in for
 * the simple gandiva filters (column op literal, AND/OR of them) that feed the
 * join and bloom operators (fpdb-executor/src/physical/filter/FilterPOp.cpp). */
typedef enum {
  FPDB_CMP_LT = 0, FPDB_CMP_LE = 1, FPDB_CMP_GT = 2, FPDB_CMP_GE = 3, FPDB_CMP_EQ = 4, FPDB_CMP_NE = 5,
  FPDB_CMP_OR = 0x10
} fpdb_cmp_op;
fpdb_status fpdb_compare_bitmap(fpdb_context* ctx, const fpdb_column* column, int op, int64_t int_constant,
                                double float_constant, const uint8_t* d_and_bitmap, int64_t num_rows,
                                uint8_t* d_bitmap, int64_t* d_num_selected);

/* Keeps the rows whose bitmap bit is set, in order, for every column
 * (arrow::compute::Filter at BloomFilterUseKernel.cpp:195; gandiva
 * SelectionVector filtering at BloomFilterUseKernel.cpp:170-177).
 * out[i].length is the number of rows out[i] has room for; rows past the
 * smallest of them are not written, and the call then fails with
 * FPDB_ERR_CAPACITY with *num_out set to the number of rows selected.  An
 * output validity bitmap (required when the input column has one) must hold
 * ceil(num_rows/32)*4 bytes and its offset must be a multiple of 32.
 * *num_out (host) receives the number of rows written (synchronizes). */
fpdb_status fpdb_filter_compact(fpdb_context* ctx, const fpdb_column* in, int num_columns,
                                int64_t num_rows, const uint8_t* d_bitmap,
                                fpdb_column* out, int64_t* num_out);

/* fpdb_filter_compact for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Allocates nothing: the tile counts and
 * bases live in the library's static per-device scratch array (768 KiB), so
 * num_rows may be at most 268,435,456 (2^28; larger inputs fail with
 * FPDB_ERR_INVALID).  Asynchronous: the selected count is written to
 * d_num_out (device int64); rows past the smallest out[i].length are not
 * written, and the caller compares *d_num_out with it after synchronizing.
 * Every call on one GPU shares that GPU's scratch array: calls that overlap
 * on one GPU (different streams or threads) race; calls on one stream are
 * ordered and safe. */
fpdb_status fpdb_filter_compact_noalloc(fpdb_context* ctx, const fpdb_column* in, int num_columns,
                                        int64_t num_rows, const uint8_t* d_bitmap, fpdb_column* out,
                                        int64_t* d_num_out);

/* Fused blocked-bloom use: Find + compaction in one pass over the columns
 * (BloomFilterUseKernel::filterRecordBatch, BloomFilterUseKernel.cpp:180-207).
 * key_index[j] names which of `in` are the key columns.  d_bitmap may be NULL.
 * out[i].length is the rows out[i] has room for (num_rows always suffices);
 * with fewer than the selected count the call writes what fits and fails with
 * FPDB_ERR_CAPACITY, *num_out holding the selected count.  *num_out (host)
 * receives the selected row count (synchronizes). */
fpdb_status fpdb_bloom_blocked_filter(fpdb_context* ctx, const uint64_t* d_blocks, int log_num_blocks,
                                      const fpdb_column* in, int num_columns,
                                      const int* key_index, int num_keys, int64_t num_rows,
                                      uint8_t* d_bitmap, fpdb_column* out, int64_t* num_out);

/* ---- hash join ------------------------------------------------------------ */

/* Join types of fpdb::plan::prephysical::JoinType
 * (fpdb-plan/include/fpdb/plan/prephysical/JoinType.h:13-20): LEFT = build
 * side kept, RIGHT = probe side kept, *_SEMI likewise. */
typedef enum {
  FPDB_JOIN_INNER = 0, FPDB_JOIN_LEFT = 1, FPDB_JOIN_RIGHT = 2, FPDB_JOIN_FULL = 3,
  FPDB_JOIN_LEFT_SEMI = 4, FPDB_JOIN_RIGHT_SEMI = 5
} fpdb_join_type;

typedef struct fpdb_hash_table fpdb_hash_table;

/* Builds a table in HBM over the build side's key columns (one integer
 * column, or two integer columns of <= 4 bytes each).  A single key column
 * whose non-null keys are distinct and span at most twice as many values as
 * there are keys (surrogate keys) gets a dense table indexed by key - min key
 * (no hashing, one read per probe row); any other build side gets an
 * open-addressing hash table.  Deciding reads back one 32-byte record, so the
 * call synchronizes the context stream once for a single-column build.  `payload`
 * (may be NULL) is one 4-byte build column stored inline in the table slots so
 * the probe reads it together with the key.  Rows with a null key are never
 * inserted.  track_build_matches != 0 keeps one flag per build row, set when a
 * probe matches it (needed for LEFT, FULL and LEFT_SEMI joins).
 * Replaces the build phase of the acero "hashjoin" node driven by
 * HashJoinArrowKernel::joinBuildTupleSet/finalizeInput
 * (fpdb-executor/src/physical/join/hashjoin/HashJoinArrowKernel.cpp:35-93,353-372)
 * and TupleSetIndex::build (fpdb-tuple/src/TupleSetIndex.cpp:44-88). */
fpdb_status fpdb_hj_build(fpdb_context* ctx, const fpdb_column* keys, int num_keys,
                          int64_t num_rows, const fpdb_column* payload,
                          int track_build_matches, fpdb_hash_table** out);
fpdb_status fpdb_hj_table_destroy(fpdb_hash_table* table);
/* Build rows seen, rows inserted (non-null keys), table slots, and whether all
 * inserted keys are distinct (then a probe row has at most one match). */
fpdb_status fpdb_hj_table_info(fpdb_hash_table* table, int64_t* build_rows, int64_t* inserted,
                               int64_t* capacity, int* unique_keys);
/* The table's form: 0 = open-addressing hash table, 1 = dense (direct-index)
 * table.  No counterpart in the reference (Arrow's SwissTable has one form);
 * for tests and diagnostics. */
fpdb_status fpdb_hj_table_form(fpdb_hash_table* table, int* form);

/* Probes and materialises in one pass.  Output rows are in probe-row order
 * (the order among the matches of one probe row is unspecified).  Output
 * columns are the probe columns listed in probe_out (row-copied) followed by
 * the build columns listed in build_out (gathered by build row;
 * build_out[j].values == NULL selects the inline payload).  By join_type:
 *   INNER       matched pairs;
 *   RIGHT       matched pairs, plus unmatched probe rows with null build columns;
 *   RIGHT_SEMI  probe rows with at least one match (build_out ignored);
 *   LEFT, FULL  as INNER, RIGHT, and matched build rows are flagged; after the
 *               last probe, fpdb_hj_build_side_rows(matched = 0) lists the
 *               unmatched build rows to append with null probe columns;
 *   LEFT_SEMI   writes nothing, flags matched build rows; afterwards
 *               fpdb_hj_build_side_rows(matched = 1) lists the result rows.
 * out[] needs room for out_capacity rows; output validity bitmaps (required
 * for nullable results) need ceil(out_capacity/32)*4 bytes.  If more rows
 * would be produced the call fails with FPDB_ERR_CAPACITY and *num_out holds
 * the required count.  *num_out (host) receives the row count (synchronizes).
 * Replaces the probe + materialisation of the acero "hashjoin" node
 * (HashJoinArrowKernel.cpp:126-206,305-351) and RecordBatchHashJoiner::join
 * (fpdb-executor/src/physical/join/hashjoin/RecordBatchHashJoiner.cpp:38-131). */
fpdb_status fpdb_hj_probe(fpdb_context* ctx, fpdb_hash_table* table, int join_type,
                          const fpdb_column* probe_keys, int num_keys, int64_t num_rows,
                          const fpdb_column* probe_out, int num_probe_out,
                          const fpdb_column* build_out, int num_build_out,
                          fpdb_column* out, int64_t out_capacity, int64_t* num_out);

/* fpdb_hj_probe without waiting: the produced row count is written to
 * d_num_out (device int64) by the kernel; rows past out_capacity are counted
 * but not written.  The caller checks *d_num_out <= out_capacity once the
 * stream has been synchronized. */
fpdb_status fpdb_hj_probe_async(fpdb_context* ctx, fpdb_hash_table* table, int join_type,
                                const fpdb_column* probe_keys, int num_keys, int64_t num_rows,
                                const fpdb_column* probe_out, int num_probe_out,
                                const fpdb_column* build_out, int num_build_out,
                                fpdb_column* out, int64_t out_capacity, int64_t* d_num_out);

/* The number of rows fpdb_hj_probe would produce (synchronizes). */
fpdb_status fpdb_hj_probe_count(fpdb_context* ctx, fpdb_hash_table* table, int join_type,
                                const fpdb_column* probe_keys, int num_keys, int64_t num_rows,
                                int64_t* count);

/* After all probes: build rows never matched (matched = 0) or matched at least
 * once (matched = 1), as int64 row indices in build order (the acero
 * RIGHT_OUTER / RIGHT_SEMI sides that LEFT / LEFT_SEMI map to,
 * HashJoinArrowKernel.cpp:165-195).  *num_out (host) receives the count. */
fpdb_status fpdb_hj_build_side_rows(fpdb_context* ctx, fpdb_hash_table* table, int matched,
                                    int64_t* d_rows, int64_t capacity, int64_t* num_out);

/* Copies num_rows rows of each input column into the output columns starting
 * at output row out_row, validity bits included at any bit offset (the device
 * side of TupleSet::append / concatenate, fpdb-tuple/src/TupleSet.cpp). */
fpdb_status fpdb_copy_rows(fpdb_context* ctx, const fpdb_column* in, int num_columns, int64_t num_rows,
                           fpdb_column* out, int64_t out_row);

/* Gathers rows (int64 indices; -1 = null row) of columns into out (used to
 * materialise build-side rows and the null probe side of outer joins). */
fpdb_status fpdb_gather(fpdb_context* ctx, const fpdb_column* in, int num_columns,
                        const int64_t* d_rows, int64_t num_rows, fpdb_column* out);

/* ---- utf8 columns and composite keys ---------------------------------------- */

/* The bytes fpdb_utf8_take would write: the sum of the lengths of the rows
 * named by `rows` (an INT64 or INT32 column; a null or negative entry names
 * no row).  *bytes (host) receives it (synchronizes). */
fpdb_status fpdb_utf8_take_bytes(fpdb_context* ctx, const fpdb_column* in, const fpdb_column* rows,
                                 int64_t num_rows, int64_t* bytes);

/* Gathers rows of a utf8 column: out row i is in row rows[i], or null where
 * rows[i] is null or negative (out then needs a validity bitmap of
 * ceil(num_rows/32)*4 bytes; it is needed too when `in` has one).  out->values
 * needs num_rows + 1 int32 offsets (out->offset must be 0) and out->data room
 * for the bytes (out->data_bytes; fpdb_utf8_take_bytes gives the count, more
 * is an FPDB_ERR_CAPACITY failure before anything is written; synchronizes).
 * The take/append of string arrays the reference leaves to Arrow
 * (arrow::compute::Take in the acero hash join's materialisation,
 * HashJoinArrowKernel.cpp:126-206; ArrayAppender in RecordBatchHashJoiner.cpp:83-119;
 * arrow::ConcatenateTables in TupleSet::append, fpdb-tuple/src/TupleSet.cpp). */
fpdb_status fpdb_utf8_take(fpdb_context* ctx, const fpdb_column* in, const fpdb_column* rows,
                           int64_t num_rows, fpdb_column* out);

/* Writes start, start + 1, ... into n int64 values (row numbers to carry
 * through a join or compaction and resolve utf8 columns with afterwards). */
fpdb_status fpdb_iota_int64(fpdb_context* ctx, int64_t* d_out, int64_t n, int64_t start);

/* Comparison of a utf8 column with a constant string (bytewise, as unsigned
 * bytes, a prefix before any longer string: Arrow's binary ordering) into a
 * selection bitmap; the other arguments and the result as for
 * fpdb_compare_bitmap (gandiva's string comparisons in FilterPOp,
 * fpdb-executor/src/physical/filter/FilterPOp.cpp). */
fpdb_status fpdb_compare_bitmap_utf8(fpdb_context* ctx, const fpdb_column* column, int op,
                                     const char* constant, int64_t constant_len,
                                     const uint8_t* d_and_bitmap, int64_t num_rows, uint8_t* d_bitmap,
                                     int64_t* d_num_selected);

/* A dictionary of key tuples in HBM: it gives each distinct row of its key
 * columns (integer, float, date and utf8 columns in any number, up to 8) a
 * dense int32 id, in order of first occurrence (a tuple whose 64-bit hash
 * equals another's may take its id a little later).  Equality is exact:
 * tuples are compared value by value, integers by value whatever their width,
 * floats by value (0.0 == -0.0), strings byte by byte.  It lets the join and
 * group-by kernels, which take one integer key, work on keys of any width and
 * type: the role of Arrow's Grouper in the acero aggregate node
 * (GroupArrowKernel.cpp:245-291) and of the key encoder of the acero hash
 * join (HashJoinArrowKernel.cpp:138-145), both in the absent Arrow library.
 * hash_bits < 64 keeps only that many bits of each tuple's hash (tests use it
 * to make distinct tuples share hashes); pass 64. */
typedef struct fpdb_keydict fpdb_keydict;
fpdb_status fpdb_keydict_create(fpdb_context* ctx, const int32_t* key_types, int num_keys, int hash_bits,
                                fpdb_keydict** out);
fpdb_status fpdb_keydict_destroy(fpdb_keydict* dict);
fpdb_status fpdb_keydict_size(fpdb_keydict* dict, int64_t* num_ids);

/* Maps num_rows rows of the key columns (same count as the dictionary's; an
 * integer column may differ in width from the dictionary's but not in
 * signedness) to ids in `ids` (an INT32 output column).  insert != 0 adds the
 * tuples not yet present (new ids); insert == 0 gives -1 for them.
 * null_is_key != 0 makes null a value like any other (a group-by key);
 * null_is_key == 0 gives -1 to a row with a null key column (a join key never
 * matches).  Where an id is -1 the ids column's validity bit (if it has a
 * bitmap) is cleared.  *num_ids (host, may be NULL) receives the dictionary's
 * size afterwards (synchronizes). */
fpdb_status fpdb_keydict_encode(fpdb_context* ctx, fpdb_keydict* dict, const fpdb_column* keys, int num_keys,
                                int64_t num_rows, int insert, int null_is_key, fpdb_column* ids,
                                int64_t* num_ids);

/* The values of key column `key` of the tuples with the given ids (an INT32
 * or INT64 column; null or negative gives null): a column of the
 * dictionary's type for that key.  A utf8 output follows fpdb_utf8_take;
 * fpdb_keydict_decode_bytes gives its byte count (synchronizes). */
fpdb_status fpdb_keydict_decode_bytes(fpdb_context* ctx, fpdb_keydict* dict, int key, const fpdb_column* ids,
                                      int64_t num_rows, int64_t* bytes);
fpdb_status fpdb_keydict_decode(fpdb_context* ctx, fpdb_keydict* dict, int key, const fpdb_column* ids,
                                int64_t num_rows, fpdb_column* out);

/* ---- wire forms (filters and bitmaps shipped to the store) --------------------- */

/* ORs an n-row selection bitmap (LSB-first bytes, as produced by the find and
 * compare entry points) into int64 bitmap words at row row_offset, bit r % 64
 * of word r / 64 being row r: FilterPOp::bufferBitMap
 * (fpdb-executor/src/physical/filter/FilterPOp.cpp:264-279) with
 * fpdb::util::setBit (fpdb-util/src/Util.cpp:115-117).  d_words must hold
 * ceil((row_offset + n) / 64) words; they are not cleared (batch after batch
 * at a growing row_offset builds the whole bitmap).  The words are the int64
 * "bitmap" column of ArrowSerializer::bitmap_to_recordBatch
 * (fpdb-tuple/src/serialization/ArrowSerializer.cpp:330-348). */
fpdb_status fpdb_bitmap_to_words(fpdb_context* ctx, const uint8_t* d_bitmap, int64_t n, int64_t row_offset,
                                 int64_t* d_words);
/* Rows [row_offset, row_offset + n) of int64 bitmap words as a selection
 * bitmap ((n + 7) / 8 bytes): FilterPOp::makeSelectionVector
 * (FilterPOp.cpp:281-297) with fpdb::util::getBit (Util.cpp:123-125). */
fpdb_status fpdb_bitmap_from_words(fpdb_context* ctx, const int64_t* d_words, int64_t num_words,
                                   int64_t row_offset, int64_t n, uint8_t* d_bitmap);
/* The two record batches of ArrowBloomFilter::makeBitmapRecordBatches
 * (fpdb-executor/src/physical/bloomfilter/ArrowBloomFilter.cpp:97-116), each
 * one uint8 column: the 136 mask bytes, then the 8 << log_num_blocks block
 * bytes (little-endian uint64 blocks).  Either output may be NULL; copying the
 * blocks synchronizes. */
fpdb_status fpdb_bloom_blocked_export(fpdb_context* ctx, const uint64_t* d_blocks, int log_num_blocks,
                                      uint8_t* masks_out, uint8_t* blocks_out);
/* ArrowBloomFilter::saveBitmapRecordBatches (ArrowBloomFilter.cpp:77-95): the
 * masks must equal BloomFilterMasks() (this library's find uses that table;
 * anything else is FPDB_ERR_INVALID); block_bytes must be 8 x a power of two.
 * Sets *log_num_blocks and, if d_blocks is not NULL, uploads the blocks. */
fpdb_status fpdb_bloom_blocked_import(fpdb_context* ctx, const uint8_t* masks, const uint8_t* blocks,
                                      uint64_t block_bytes, int* log_num_blocks, uint64_t* d_blocks);

/* ---- device-resident segment cache ---------------------------------------------- */

/* Column segments of table partitions kept in HBM under (partition, column)
 * keys -- the range is carried but, as in the reference, not part of a key's
 * identity (fpdb-cache/src/SegmentKey.cpp:58-72) -- with a size-bounded
 * replacement policy.  Segment size = values bytes + ceil(length/8) validity
 * bytes (the Arrow buffer sizes the reference's SegmentMetadata counts).
 * Host bookkeeping, device data: SegmentCache (fpdb-cache/src/SegmentCache.cpp)
 * with LRUCachingPolicy (fpdb-cache/src/policy/LRUCachingPolicy.cpp). */
typedef struct fpdb_segment_cache fpdb_segment_cache;
typedef enum { FPDB_CACHE_LRU = 0 } fpdb_cache_policy;
fpdb_status fpdb_cache_create(fpdb_context* ctx, uint64_t capacity_bytes, int policy, fpdb_segment_cache** out);
fpdb_status fpdb_cache_destroy(fpdb_segment_cache* cache);
/* SegmentCache::store (SegmentCache.cpp:21-30): copies `data` (host or device
 * memory) into cache-owned HBM, evicting least recently used segments until
 * it fits.  *stored = 0 when the segment is larger than the whole cache
 * (LRUCachingPolicy::onStore returns no decision, LRUCachingPolicy.cpp:32-38).
 * Storing a key that is present refreshes it. */
fpdb_status fpdb_cache_store(fpdb_segment_cache* cache, const char* partition, const char* column,
                             int64_t start_offset, int64_t finish_offset, const fpdb_column* data, int* stored);
/* SegmentCache::load (SegmentCache.cpp:32-45): on a hit *hit = 1 and *out
 * describes the cached device buffers, pinned until fpdb_cache_release (an
 * eviction meanwhile drops the key but keeps the memory).  A load moves the
 * key to the front of the usage queue. */
fpdb_status fpdb_cache_load(fpdb_segment_cache* cache, const char* partition, const char* column,
                            int64_t start_offset, int64_t finish_offset, fpdb_column* out, int* hit);
fpdb_status fpdb_cache_release(fpdb_segment_cache* cache, const void* values);
/* SegmentCache::remove(predicate) (SegmentCache.cpp:51-61) for one partition,
 * or one column of it (column may be NULL). */
fpdb_status fpdb_cache_remove(fpdb_segment_cache* cache, const char* partition, const char* column,
                              int64_t* removed);
/* SegmentCache::clear (SegmentCache.cpp:80-84): every segment and the counters. */
fpdb_status fpdb_cache_clear(fpdb_segment_cache* cache);
/* Segments held, policy bytes used / free, and load hits / misses. */
fpdb_status fpdb_cache_info(fpdb_segment_cache* cache, int64_t* num_segments, uint64_t* used_bytes,
                            uint64_t* free_bytes, int64_t* hits, int64_t* misses);
/* The usage queue, most recent first, as "partition\tcolumn\n" lines
 * (LRUCachingPolicy::showCurrentLayout, LRUCachingPolicy.cpp:79-85);
 * *needed = bytes including the terminating NUL. */
fpdb_status fpdb_cache_layout(fpdb_segment_cache* cache, char* buf, uint64_t buf_len, uint64_t* needed);

/* ---- hash shuffle ------------------------------------------------------------ */

/* Splits num_rows rows into num_slots partitions: row r goes to partition
 * hash(r) % num_slots, hash being the Arrow 6 row hash of the key columns (as
 * fpdb_hash_keys computes it), and each partition keeps the input row order.
 * Every column of `in` is written to the matching column of `out` (same
 * width, room for num_rows rows) partition after partition, so partition p
 * holds output rows [offsets[p], offsets[p + 1]).  The offsets (num_slots + 1
 * int64) go to d_offsets (device, may be NULL) and to offsets (host, may be
 * NULL; non-NULL synchronizes).  Output validity bitmaps (required for
 * nullable inputs) need ceil(num_rows/32)*4 bytes and an offset that is a
 * multiple of 32.  1 <= num_slots <= 4096.  Replaces ShuffleKernel2::shuffle
 * (fpdb-executor/src/physical/shuffle/ShuffleKernel2.cpp:28-126: partition =
 * hash % numSlots at :68, partitions concatenated in slot order at :85-97 and
 * sliced into one TupleSet per slot at :113-119), which ShufflePOp::onTupleSet
 * calls with numSlots = the number of consumers
 * (fpdb-executor/src/physical/shuffle/ShufflePOp.cpp:135-170). */
fpdb_status fpdb_shuffle(fpdb_context* ctx, const fpdb_column* keys, int num_keys,
                         const fpdb_column* in, int num_columns, int64_t num_rows, int num_slots,
                         fpdb_column* out, int64_t* d_offsets, int64_t* offsets);

/* ---- hash group-by --------------------------------------------------------- */

/* Aggregate kinds of fpdb::executor::physical::aggregate
 * (Sum.cpp:63-74 hash_sum; Count.cpp:91-104 hash_count ALL / ONLY_VALID;
 * MinMax.cpp:84 hash_min / hash_max). */
typedef enum {
  FPDB_AGG_COUNT_ALL = 0, FPDB_AGG_COUNT_VALID = 1, FPDB_AGG_SUM = 2,
  FPDB_AGG_MIN = 3, FPDB_AGG_MAX = 4,
  /* hash_sum / hash_count then "divide" and a cast to float64 (AvgBase::finalize,
   * fpdb-executor/src/physical/aggregate/function/AvgBase.cpp:79-96): integer
   * sums divide by the count with truncation, float sums divide exactly. */
  FPDB_AGG_MEAN = 5
} fpdb_agg_kind;

typedef struct fpdb_agg_spec {
  int32_t kind;             /* fpdb_agg_kind                                  */
  int32_t input_type;       /* fpdb_type of the input column (COUNT_ALL: any) */
  int32_t input_nullable;   /* input batches may carry a validity bitmap      */
  int32_t reserved;
} fpdb_agg_spec;

typedef struct fpdb_group_table fpdb_group_table;

/* Creates an accumulating grouped aggregation over key columns of the given
 * types (none: one group for all rows; one of <= 8 bytes; or two of <= 4
 * bytes; no nulls) for at most max_groups groups.  Replaces GroupArrowKernel's
 * acero "aggregate" node
 * (fpdb-executor/src/physical/group/GroupArrowKernel.cpp:16-68,127-154,245-291). */
fpdb_status fpdb_groupby_create(fpdb_context* ctx, const int32_t* key_types, int num_keys,
                                const fpdb_agg_spec* aggs, int num_aggs, int64_t max_groups,
                                fpdb_group_table** out);

/* Same, for one to four integer key columns packed into one 64-bit key: key j
 * holds values in [key_min[j], key_min[j] + 2^key_bits[j]), the bits summing
 * to at most 64.  A value outside its range makes fpdb_groupby_num_groups and
 * fpdb_groupby_finalize fail with FPDB_ERR_INVALID.  The grouping of the
 * reference's multi-column group keys (GroupArrowKernel.cpp:245-291, any
 * number of keys) for key ranges that fit; fpdb_column_min_max gives them. */
fpdb_status fpdb_groupby_create_packed(fpdb_context* ctx, const int32_t* key_types, int num_keys,
                                       const int64_t* key_min, const int32_t* key_bits,
                                       const fpdb_agg_spec* aggs, int num_aggs, int64_t max_groups,
                                       fpdb_group_table** out);
fpdb_status fpdb_groupby_destroy(fpdb_group_table* table);

/* Smallest and largest valid value of an integer column (0 and 0 when it has
 * none) and the number of valid values (may be NULL); synchronizes. */
fpdb_status fpdb_column_min_max(fpdb_context* ctx, const fpdb_column* column, int64_t num_rows,
                                int64_t* min, int64_t* max, int64_t* num_valid);

/* Consumes one batch (GroupArrowKernel::group / doGroup, GroupArrowKernel.cpp:16-39,127-154).
 * values[j] is the input of aggs[j] (ignored for COUNT_ALL). */
fpdb_status fpdb_groupby_consume(fpdb_context* ctx, fpdb_group_table* table,
                                 const fpdb_column* keys, const fpdb_column* values,
                                 int64_t num_rows);

/* Number of groups so far (synchronizes). */
fpdb_status fpdb_groupby_num_groups(fpdb_context* ctx, fpdb_group_table* table, int64_t* num_groups);

/* Output type of an aggregate: COUNT -> INT64; SUM of signed integers ->
 * INT64, of unsigned -> UINT64, of floats -> FLOAT64; MIN/MAX -> input type;
 * MEAN -> FLOAT64. */
int32_t fpdb_groupby_output_type(fpdb_agg_spec spec);

/* Writes the groups in first-appearance order (the serial grouper's group-id
 * order): key columns then one column per aggregate (COUNT: int64; SUM of
 * integers: int64; SUM of floats: float64; MIN/MAX: input type).  A SUM, MIN
 * or MAX with no valid input in a group is null (ScalarAggregateOptions
 * defaults, min_count = 1).  out_keys/out_aggs need room for num_groups rows
 * (GroupArrowKernel::finalise, GroupArrowKernel.cpp:41-68). */
fpdb_status fpdb_groupby_finalize(fpdb_context* ctx, fpdb_group_table* table,
                                  fpdb_column* out_keys, fpdb_column* out_aggs,
                                  int64_t* num_groups);

/* Row-wise num / den as float64: an average from its intermediate sum and
 * count columns, as AvgBase::finalize computes it for Avg and AvgReduce
 * (fpdb-executor/src/physical/aggregate/function/AvgBase.cpp:71-98: arrow
 * "divide", then a cast to float64).  Two integer columns divide as integers
 * (truncating toward zero) before the cast; otherwise the quotient is taken in
 * float64.  A null num or den gives a null; an integer division by zero fails
 * with FPDB_ERR_INVALID, as arrow's integer "divide" does.  out is float64
 * with a validity bitmap of ceil(n/32)*4 bytes at an offset that is a multiple
 * of 32.  Synchronizes. */
fpdb_status fpdb_divide_to_float64(fpdb_context* ctx, const fpdb_column* num, const fpdb_column* den,
                                   int64_t num_rows, fpdb_column* out);

/* ---- batched star join with predicate transfer ------------------------------ */

/* One dimension of a star join: dimension keys (non-null int64), a filter
 * `filter_attr <filter_op> filter_constant` on an int32 attribute, an optional
 * int32 group attribute, and the fact table's foreign-key column for it. */
typedef struct fpdb_star_dimension {
  fpdb_column key;
  fpdb_column filter_attr;
  int32_t filter_op;            /* fpdb_cmp_op */
  int32_t has_group_attr;
  int64_t filter_constant;
  fpdb_column group_attr;
  fpdb_column fact_fk;
} fpdb_star_dimension;

/* SELECT g_a, g_b, COUNT(*), SUM(measure) FROM fact JOIN dims ... GROUP BY g_a, g_b
 * with full predicate transfer: forward blocked bloom filters from every
 * filtered dimension into the fact table, backward filters from the surviving
 * fact rows into every dimension, then the joins and the group-by; the plan
 * shape SmallToLargePredTransOrder builds
 * (fpdb-executor/src/physical/transform/pred-trans/SmallToLargePredTransOrder.cpp:38-214),
 * over a fact table split into partitions (the reference supports one
 * partition, :45-48): forward filters cover all of a dimension's kept rows,
 * backward filters the union of all partitions' survivors, and groups are
 * numbered by first appearance in fact order. */
typedef struct fpdb_star_query {
  int32_t num_dims;                   /* 1..8 */
  int32_t group_dims[2];              /* dimensions whose group_attr forms the key */
  int32_t filter_order[8];            /* order of the forward filters on the fact rows */
  fpdb_star_dimension dims[8];
  fpdb_column fact_measure;           /* non-null int64 */
  int64_t fact_rows;
  int32_t num_partitions;
  int32_t reserved;
  const int64_t* partition_offsets;   /* host, num_partitions + 1 row offsets */
  int64_t max_groups;
  /* Optional: partitions held in buffers of their own (the reference's one
   * TupleSet per partition).  When partition_fact_fk is not NULL it holds
   * num_partitions x num_dims foreign-key columns (partition p's column for
   * dimension d at p * num_dims + d) and partition_measure num_partitions
   * measure columns, each with at least its partition's rows
   * (partition_offsets[p + 1] - partition_offsets[p]); dims[d].fact_fk and
   * fact_measure are then not read.  Rows keep their global numbers (offset
   * + row in partition) for the first-appearance order of the groups. */
  const fpdb_column* partition_fact_fk;
  const fpdb_column* partition_measure;
} fpdb_star_query;

/* Launch shape of the batched form (0 = default): workgroups per CU of the
 * persistent kernel, hash-table capacity as a power-of-two multiple
 * (2^ht_load_shift) of the dimension rows, and fact rows per thread per step
 * of its fact pass (4, 8 or 16; 0 = 8).  Chosen by the autotuner (with the
 * query's filter_order).  barrier_timeout_us bounds each grid-barrier wait of
 * the persistent kernel (0 = 200 ms); a wait past it ends the launch early
 * and the result call reports FPDB_ERR_INTERNAL. */
typedef struct fpdb_star_tuning {
  int32_t blocks_per_cu;
  int32_t barrier_timeout_us;
  int32_t ht_load_shift;
  int32_t items_per_thread;
} fpdb_star_tuning;

typedef struct fpdb_star fpdb_star;

/* Allocates every buffer of both executions up front (no allocation while
 * running, so the batched form can be captured in a graph).  No counterpart in
 * the reference, which plans the same chain as separate operators. */
fpdb_status fpdb_star_prepare(fpdb_context* ctx, const fpdb_star_query* query, const fpdb_star_tuning* tuning,
                              fpdb_star** out);
fpdb_status fpdb_star_destroy(fpdb_star* star);
/* The plain form: per-partition launches of the operator entry points
 * (fpdb_bloom_blocked_filter, fpdb_bloom_blocked_insert, fpdb_hj_build,
 * fpdb_hj_probe, fpdb_groupby_*), synchronizing as the operators do. */
fpdb_status fpdb_star_run_plain(fpdb_context* ctx, fpdb_star* star);
/* The batched form: one persistent launch over all partitions (asynchronous). */
fpdb_status fpdb_star_run_batched(fpdb_context* ctx, fpdb_star* star);
/* Captures the batched form in a hipGraph once; launches replay it. */
fpdb_status fpdb_star_graph_capture(fpdb_context* ctx, fpdb_star* star);
fpdb_status fpdb_star_graph_launch(fpdb_context* ctx, fpdb_star* star);
/* Result of the last run (host arrays with room for `capacity` groups). */
fpdb_status fpdb_star_result(fpdb_context* ctx, fpdb_star* star, int32_t* g_a, int32_t* g_b, int64_t* count,
                             int64_t* sum, int64_t capacity, int64_t* num_groups);
/* A forward (backward = 0) or backward filter of the last run, for parity
 * checks: its blocks (if `blocks` is not NULL), log2 block count and the
 * number of input rows it was sized for. */
fpdb_status fpdb_star_filter(fpdb_context* ctx, fpdb_star* star, int dim, int backward, uint64_t* blocks,
                             int64_t capacity, int* log_num_blocks, int64_t* filter_input_rows);

/* ---- serving loop: continuous batching over a resident star schema ------- */

/* A dimension of the resident schema: dense keys (key[i] == i), an int32
 * filter attribute, an optional int32 group attribute with values in
 * [0, group_domain), and the fact table's foreign keys into it. */
typedef struct fpdb_serve_dimension {
  fpdb_column key;
  fpdb_column filter_attr;
  fpdb_column group_attr;     /* values NULL: no group attribute */
  int32_t group_domain;
  int32_t reserved;
  fpdb_column fact_fk;
} fpdb_serve_dimension;

typedef struct fpdb_serve_schema {
  int32_t num_dims;           /* 1..8 */
  int32_t reserved;
  fpdb_serve_dimension dims[8];
  fpdb_column fact_measure;   /* int64, non-null */
  int64_t fact_rows;
} fpdb_serve_schema;

typedef struct fpdb_serve_config {
  int32_t max_active;         /* queries sharing the scan at once, 1..8 */
  int32_t chunks;             /* chunks per pass over the fact table (0: 8) */
  uint64_t workspace_bytes;   /* HBM budget for the queries in the scan (0: none) */
} fpdb_serve_config;

/* SELECT g0, g1, COUNT(*), SUM(measure) FROM fact JOIN dims WHERE <filters>
 * GROUP BY <group attributes> - the star join + group-by the predicate-transfer
 * plans of SmallToLargePredTransOrder.cpp:38-214 produce, for a serving
 * workload (no counterpart in the reference, which runs one query plan at a
 * time).  Filters on one dimension are AND-ed. */
typedef struct fpdb_serve_query {
  int32_t num_filters;        /* 0..4 */
  int32_t num_group;          /* 0..2 (0: one group over all passing rows) */
  int32_t filter_dim[4];
  int32_t filter_op[4];       /* fpdb_cmp_op */
  int64_t filter_constant[4];
  int32_t group_dim[2];
} fpdb_serve_query;

typedef struct fpdb_serve fpdb_serve;

/* Registers the resident schema (borrowed: the columns must outlive the
 * server) and checks that dimension keys are dense.  The server keeps a 4-byte
 * copy of each fact foreign-key column (4 B x fact rows per dimension of HBM,
 * counted against the context's budget) for the scan to read. */
fpdb_status fpdb_serve_create(fpdb_context* ctx, const fpdb_serve_schema* schema,
                              const fpdb_serve_config* config, fpdb_serve** out);
fpdb_status fpdb_serve_destroy(fpdb_serve* server);
/* Queues a query (thread-safe, and never waits for a step in progress: the
 * step holds the server's lock only while it admits and books queries);
 * *query_id identifies it from then on. */
fpdb_status fpdb_serve_submit(fpdb_serve* server, const fpdb_serve_query* query, int64_t* query_id);
/* One tick of the circular shared scan.  The fact table is read in `chunks`
 * chunks, one per tick, round and round; a query is admitted at any tick
 * (FIFO, within max_active, 8192 groups and the workspace budget), joins the
 * scan at the current chunk and is finished after one full pass, however
 * many other queries come and go meanwhile.  Its finalize (partials added in
 * a fixed order, groups sorted by first appearance) and the copy of its
 * groups to the host run on a second stream, overlapped with the next ticks;
 * a step collects the queries whose copies have landed.  A query whose
 * workspace cannot be allocated at admission fails on its own (state 3,
 * fpdb_serve_result returns the error) without holding up the others.
 * *completed = queries finished or failed during this call.  With nothing
 * queued or scanning, a step waits for the finalizes still in flight. */
fpdb_status fpdb_serve_step(fpdb_serve* server, int* completed);
/* 0 queued, 1 in the scan (or being finalized), 2 finished, 3 failed. */
fpdb_status fpdb_serve_status(fpdb_serve* server, int64_t query_id, int* state);
/* A finished query's groups in first-appearance order (host arrays; g1 is 0
 * for one group dimension, g0 and g1 are 0 for none); for a failed query,
 * the error that failed it. */
fpdb_status fpdb_serve_result(fpdb_serve* server, int64_t query_id, int32_t* g0, int32_t* g1, int64_t* count,
                              int64_t* sum, int64_t capacity, int64_t* num_groups);
/* Replays an arrival trace in real time (arrival_ns non-decreasing, relative
 * to the call): submits each query at its arrival, steps while work is queued
 * or in flight, and reports each query's id and latency (the moment its
 * result is on the host, minus its arrival; -1 for a failed query) and the
 * whole trace's makespan. */
fpdb_status fpdb_serve_run_trace(fpdb_serve* server, const fpdb_serve_query* queries, const int64_t* arrival_ns,
                                 int64_t num_queries, int64_t* query_ids, int64_t* latency_ns,
                                 int64_t* makespan_ns);

#ifdef __cplusplus
}
#endif

#endif /* FPDB_AMD_H */
