/*
 * q1_hip.h -- C ABI of the MI355X (gfx950) TPC-H Q1 scan -> filter -> group-aggregate path.
 *
 * Plain C: POD structs, plain pointers and sizes, status codes; no exception crosses the
 * boundary and no torch or HIP type appears in a signature (streams are passed as void*).
 *
 * Reference interfaces replaced (diegomestre2/tpchQ01_GPU):
 *   - the kernel maps `kernels`, `kernels_compressed`, `kernels_filter_pushdown`
 *     (src/main.cu:166-191) and their launch in execute_query_1_once (src/main.cu:657-718)
 *     -> q1_enqueue (same accumulate-into-zeroed-aggregates contract, src/main.cu:553-558);
 *   - the whole of execute_query_1_once (src/main.cu:498-736) on device-resident columns
 *     -> q1_run; on host-resident columns streamed in chunks -> q1_run_streaming;
 *   - the CPU/GPU split (src/main.cu:513-550, src/cpu.h:23-38) -> q1_run_split / q1_cpu_run;
 *   - the filter-pushdown producer precompute_filter_for_table_chunk (src/main.cu:362-389)
 *     -> q1_build_filter (device) / q1_cpu_build_filter (host);
 *   - the column compression loop in main (src/main.cu:773-803) -> q1_compress_columns;
 *   - `--device` (src/helper.hpp:18-48) -> q1_device_info.
 * Entry points with no reference counterpart are marked "(new)".
 */
#ifndef Q1_HIP_H_
#define Q1_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define Q1_ABI_VERSION 2  /* 2: q1_server_config gained exclusive_bytes */
#define Q1_NUM_GROUPS 6 /* slot = return_flag_code << 1 | line_status_code (src/data_types.h:52-80) */

typedef enum {
  Q1_OK = 0,
  Q1_EINVAL = 1,    /* bad argument: null pointer, unknown enum, misaligned column, range */
  Q1_ENOMEM = 2,    /* device or host allocation failed / admission refused */
  Q1_EDEVICE = 3,   /* a HIP runtime call failed */
  Q1_EOVERFLOW = 4, /* a result would not be exact in the result types */
  Q1_EBUSY = 5,     /* serving queue full */
  Q1_ETIMEOUT = 6,  /* wait timed out */
  Q1_ECANCELED = 7  /* request canceled */
} q1_status;

/* Aggregation-table placement; the reference's --hash-table-placement keys
 * (src/main.cu:166-173) in order. */
typedef enum {
  Q1_PLACE_IN_REGISTERS = 0,            /* "in_registers" / "in-registers" */
  Q1_PLACE_IN_REGISTERS_PER_THREAD = 1, /* "in_registers_per_thread" (the reference default) */
  Q1_PLACE_LOCAL_MEM = 2,               /* "local_mem" / "local-mem" */
  Q1_PLACE_PER_THREAD_SHARED_MEM = 3,   /* "shared_mem_per_thread" / "per-thread-shared-mem" */
  Q1_PLACE_GLOBAL = 4,                  /* "global" */
  Q1_PLACE_AUTO = 5                     /* (new) fastest placement per the autotuning table */
} q1_placement;

/* Column layouts.  Row i of every column describes lineitem row i. */
typedef enum {
  /* 38 B/row, the reference's lineitem columns (src/data_types.h:19-25):
   * int32 ship_date (MonetDB day), int64 quantity (x100), int64 extended_price (cents),
   * int64 discount (x100), int64 tax (x100), char return_flag, char line_status */
  Q1_LAYOUT_REFERENCE = 0,
  /* 20 B/row: int32 ship_date, int32 quantity (x100), int64 extended_price,
   * int8 discount, int8 tax, char return_flag, char line_status */
  Q1_LAYOUT_UNCOMPRESSED = 1,
  /* 9.375 B/row (src/main.cu:779-788): uint16 ship_date - 727563, uint8 quantity (x1),
   * uint32 extended_price, uint8 discount, uint8 tax, return_flag 2 bits and
   * line_status 1 bit packed LSB-first into uint32 words (16 resp. 32 rows per word) */
  Q1_LAYOUT_COMPRESSED = 2,
  /* 7.5 B/row: as COMPRESSED but the ship_date column is replaced by a precomputed
   * filter bitmap (1 bit/row, LSB-first uint32 words) for one cutoff */
  Q1_LAYOUT_COMPRESSED_PUSHDOWN = 3
} q1_layout;

/* Column pointers; element types per layout as above.  Unused pointers may be NULL. */
typedef struct {
  const void* ship_date;
  const void* quantity;
  const void* extended_price;
  const void* discount;
  const void* tax;
  const void* return_flag;
  const void* line_status;
  const void* filter; /* COMPRESSED_PUSHDOWN only */
} q1_columns;

/* One group of the result, in the reference CPU kernels' scales (cpu/common.hpp:158-165):
 * sum_qty and sum_disc x100, sum_base_price in cents, sum_disc_price scale 1e-4 and
 * sum_charge scale 1e-6 as two's-complement 128-bit integers (lo, hi). */
typedef struct {
  int64_t sum_qty;
  int64_t sum_base_price;
  uint64_t sum_disc_price_lo;
  int64_t sum_disc_price_hi;
  uint64_t sum_charge_lo;
  int64_t sum_charge_hi;
  int64_t sum_disc;
  int64_t count;
} q1_group;

typedef struct {
  q1_group g[Q1_NUM_GROUPS];
} q1_result;

/* Device-side accumulation table for q1_enqueue (the reference's six aggregate arrays,
 * src/main.cu:827-836, widened): the caller zeroes it once, every launch adds into it.
 * sum_qty is always kept x100.  128-bit fields are (lo, hi) word pairs. */
typedef struct {
  uint64_t sum_qty[Q1_NUM_GROUPS];
  uint64_t sum_base_price[Q1_NUM_GROUPS];
  uint64_t sum_disc_price_lo[Q1_NUM_GROUPS];
  uint64_t sum_disc_price_hi[Q1_NUM_GROUPS];
  uint64_t sum_charge_lo[Q1_NUM_GROUPS];
  uint64_t sum_charge_hi[Q1_NUM_GROUPS];
  uint64_t sum_disc[Q1_NUM_GROUPS];
  uint64_t count[Q1_NUM_GROUPS];
} q1_device_aggregates;

/* Launch shape (the reference's --threads-per-block / --tuples-per-thread,
 * src/main.cu:255-360, 627-650).  Zero fields mean "from the autotuning table"; launches
 * of q1_enqueue and q1_run_streaming without Q1_LAUNCH_ASSUME_TPCH_DOMAIN run the
 * any-value kernel, whose zero fields mean 256 threads and 1024 tuples per thread. */
typedef struct {
  uint32_t threads_per_block; /* 64, 128 or 256 (the batched kernel: 256) */
  uint32_t tuples_per_thread; /* multiple of 16 */
  uint32_t num_blocks;        /* 0 = ceil(rows / (threads_per_block * tuples_per_thread)) */
  uint32_t flags;             /* Q1_LAUNCH_* */
} q1_launch;

/* q1_enqueue only: the caller asserts the columns lie in the TPC-H value domain
 * (discount <= 15, tax <= 15, quantity <= 63 units, extended_price < 2^24) so the narrow
 * 32-bit kernels may run; without it the 64-bit kernels run.  Tables created through
 * this library measure their own domain. */
#define Q1_LAUNCH_ASSUME_TPCH_DOMAIN 1u

typedef struct {
  double kernel_ms;      /* device time of the last run (HIP events around the kernels) */
  double h2d_ms;         /* device time of host-to-device copies of the last call */
  double wall_ms;        /* host wall-clock of the last call */
  uint64_t bytes_read;   /* algorithmic column bytes the last run read */
  uint64_t rows;         /* rows scanned */
  uint32_t num_blocks;   /* grid of the last kernel */
  uint32_t threads_per_block;
  uint32_t tuples_per_thread;
  uint32_t placement;
} q1_timing;

typedef struct {
  char name[64];
  char arch[32];
  int32_t compute_units;
  int32_t clock_khz;
  int32_t memory_clock_khz;
  int32_t memory_bus_width;
  uint64_t total_memory;
  uint64_t lds_per_block;
  int32_t wavefront_size;
  int32_t max_threads_per_block;
} q1_device_info;

typedef struct q1_ctx q1_ctx;
typedef struct q1_table q1_table;

const char* q1_strerror(q1_status s);
int q1_abi_version(void);
const char* q1_last_error(q1_ctx* ctx); /* (new) detail text of the last failure */

/* --- context ------------------------------------------------------------------ */
q1_status q1_init(int device, q1_ctx** out);
/* Releases the context.  Tables of the context that are still alive lose their device
 * memory; their handles then serve only q1_table_free. */
void q1_shutdown(q1_ctx* ctx);
q1_status q1_device_info_get(q1_ctx* ctx, q1_device_info* out);
q1_status q1_synchronize(q1_ctx* ctx);
void* q1_stream(q1_ctx* ctx); /* (new) the context's hipStream_t */

/* --- tables (device-resident columns) ------------------------------------------ */
/* Copies host columns to device memory (pinned staging, timed in q1_last_timing.h2d_ms). */
q1_status q1_table_upload(q1_ctx* ctx, q1_layout layout, const q1_columns* host, uint64_t rows,
                          q1_table** out);
/* Wraps caller-owned device columns (16-byte aligned); nothing is copied or freed.  The
 * columns' value domain (which selects the narrow kernels) is measured once, here: a
 * caller that later writes the columns calls q1_table_refresh_domain before the next run. */
q1_status q1_table_wrap(q1_ctx* ctx, q1_layout layout, const q1_columns* device, uint64_t rows,
                        q1_table** out);
/* (new) Re-measures a table's value domain on the device (after its caller-owned columns
 * were written). */
q1_status q1_table_refresh_domain(q1_ctx* ctx, q1_table* t);
/* (new) Generates a synthetic table on the device with the library's seeded generator. */
q1_status q1_table_generate(q1_ctx* ctx, q1_layout layout, uint64_t seed, uint64_t rows,
                            q1_table** out);
q1_status q1_table_info(const q1_table* t, q1_layout* layout, uint64_t* rows, q1_columns* device);
/* (new) Copies the device columns back into caller host buffers of the layout's sizes. */
q1_status q1_table_download(q1_ctx* ctx, const q1_table* t, const q1_columns* host_out);
uint64_t q1_table_device_bytes(const q1_table* t);
/* Frees the table's device memory (if its context is still alive) and the handle. */
void q1_table_free(q1_table* t);
/* Pushdown producer on the device: builds a COMPRESSED_PUSHDOWN table sharing the
 * compressed table's columns plus a filter bitmap for `cutoff_day`.  The pushdown table
 * owns only its bitmap: free it before the compressed table, and use it only while the
 * compressed table lives. */
q1_status q1_build_filter(q1_ctx* ctx, const q1_table* compressed, int32_t cutoff_day,
                          q1_table** pushdown_out);

/* --- the query ---------------------------------------------------------------- */
/* Q1 over rows [row_begin, row_begin + row_count) with `l_shipdate <= cutoff_day`
 * (ignored for PUSHDOWN tables, whose bitmap holds the filter).  Fully reduced result
 * into caller-owned host memory; no zero-first precondition.  launch may be NULL. */
q1_status q1_run(q1_ctx* ctx, const q1_table* t, uint64_t row_begin, uint64_t row_count,
                 int32_t cutoff_day, q1_placement placement, const q1_launch* launch,
                 q1_result* out);
q1_status q1_last_timing(q1_ctx* ctx, q1_timing* out);

/* The drop-in kernel launch: accumulate rows [row_begin, row_begin+num_tuples) of the
 * device columns into the caller-zeroed device table `aggregates` on `hip_stream`
 * (NULL = the context stream).  Launches on disjoint ranges compose by addition.
 * Large grids merge through spread tables: the context's on the context stream, a set
 * made on a caller stream's first large launch (kept until q1_shutdown, at most 16
 * streams; a stream under graph capture or beyond those merges directly). */
q1_status q1_enqueue(q1_ctx* ctx, q1_placement placement, q1_layout layout,
                     const q1_launch* launch, q1_device_aggregates* aggregates,
                     const q1_columns* device, uint64_t row_begin, uint64_t num_tuples,
                     int32_t cutoff_day, void* hip_stream);
/* No counterpart in the reference: q1_enqueue that allocates nothing (no hipMalloc, hipFree
 * or allocating library call), for callers that cannot allocate device memory -- code
 * under HIP graph capture, serving loops that forbid allocation.  Same arguments, results
 * and errors as q1_enqueue.  Its scratch (the spread-merge tables of grids of 176 blocks or
 * more, and the global placement's per-block tables) is a static array of 8 MiB + 24 KiB
 * per GPU, set up by q1_init and held until the process ends; a launch needing more (the
 * global placement with more than 34952 blocks) gets Q1_EINVAL.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (other streams or threads) race, so the
 * caller orders them.  The global placement may run on any stream here. */
q1_status q1_enqueue_noalloc(q1_ctx* ctx, q1_placement placement, q1_layout layout,
                             const q1_launch* launch, q1_device_aggregates* aggregates,
                             const q1_columns* device, uint64_t row_begin, uint64_t num_tuples,
                             int32_t cutoff_day, void* hip_stream);
/* Converts a host copy of the device table into a result. */
void q1_aggregates_to_result(const q1_device_aggregates* host_copy, q1_result* out);
/* Exact 128-bit merge of two results (out = a + b). */
void q1_result_merge(const q1_result* a, const q1_result* b, q1_result* out);

/* --- batched (new): many requests in one persistent, hipGraph-replayed launch ------ */
typedef struct {
  const q1_table* table;
  uint64_t row_begin;
  uint64_t row_count;
  int32_t cutoff_day;
  int32_t reserved;
} q1_request;

/* Runs n requests (at most 65536) in one launch; outs[i] is request i's result, equal to
 * q1_run of that request alone.  Requests of one layout on tables inside the TPC-H domain
 * share one launch; any other mix runs request by request.  `launch` may be NULL; its
 * threads_per_block must be 0 or 256; tuples_per_thread (rows per lane of each block of a
 * request; 0 = the tuned value) a multiple of 16, at most 256 for the shared placement and
 * 16384 for the register one; num_blocks (the resident grid, below 2^24) 0 = the tuned
 * blocks per compute unit x compute units.  The context keeps the batch's buffers between
 * calls of one capacity (the next power of two >= n, at least 16): the descriptors, the
 * results and 24 KiB of partial tables per request for the first 1024 requests. */
q1_status q1_run_batch(q1_ctx* ctx, const q1_request* reqs, uint32_t n, const q1_launch* launch,
                       q1_result* outs);

/* --- host-side helpers ----------------------------------------------------------- */
/* Allocate / free pinned host memory (hipHostMalloc). */
void* q1_host_alloc(q1_ctx* ctx, size_t bytes);
void q1_host_free(q1_ctx* ctx, void* p);
/* Host bytes per layout column for `rows` rows (packed columns in whole uint32 words). */
q1_status q1_layout_column_bytes(q1_layout layout, uint64_t rows, uint64_t bytes_out[8]);
/* (new) Fills host columns of `layout` with rows [row_begin, row_begin+rows) of the
 * seeded generator, on `threads` host threads (0 = all). */
q1_status q1_generate_host(q1_layout layout, uint64_t seed, uint64_t row_begin, uint64_t rows,
                           const q1_columns* host_out, int threads);
/* Narrows REFERENCE columns into COMPRESSED columns (src/main.cu:773-803). */
q1_status q1_compress_columns(const q1_columns* reference, uint64_t rows,
                              const q1_columns* compressed_out, int threads);
/* Host pushdown producer: filter bits for rows [0, rows) of a uint16 compressed
 * ship_date column (src/main.cu:362-389 semantics, any cutoff). */
q1_status q1_cpu_build_filter(const uint16_t* compressed_ship_date, uint64_t rows,
                              int32_t cutoff_day, uint32_t* filter_out, int threads);
/* (new) Whether host columns of `layout` lie inside the TPC-H value domain (discount and
 * tax 0..15, quantity 0..63 units, extended_price below 2^24), i.e. whether a caller may
 * pass Q1_LAUNCH_ASSUME_TPCH_DOMAIN for them; *in_domain = 1 or 0.  Tables created
 * through this library measure their own domain on the device. */
q1_status q1_host_columns_in_tpch_domain(q1_layout layout, const q1_columns* host, uint64_t rows, int threads,
                                         int* in_domain);
/* MonetDB day number of a calendar date (cpu/monetdb.cpp:61-77 numbering). */
int32_t q1_day_number(int32_t year, int32_t month, int32_t day);

/* --- CPU coprocessor (src/cpu.h:23-38, cpu/kernel_x100.hpp:422-500) ------------------ */
/* Morsel-parallel aggregation of host columns on `threads` host threads (0 = all). */
q1_status q1_cpu_run(q1_layout layout, const q1_columns* host, uint64_t row_begin,
                     uint64_t row_count, int32_t cutoff_day, int threads, q1_result* out);
/* CPU/GPU split: the device aggregates rows [0, split) of `device_table` while host
 * threads aggregate [split, rows) of `host`; split = rows - rows*cpu_fraction rounded
 * down to a multiple of `align_rows` (src/main.cu:513-529).  Exact merge. */
q1_status q1_run_split(q1_ctx* ctx, const q1_table* device_table, const q1_columns* host,
                       uint64_t rows, int32_t cutoff_day, q1_placement placement,
                       const q1_launch* launch, double cpu_fraction, uint64_t align_rows,
                       int cpu_threads, q1_result* out, double* cpu_ms, double* gpu_ms);

/* Host-resident columns streamed to the device in chunks over `num_streams` streams
 * with copies overlapping kernels (src/main.cu:567-724).  `host` should be pinned.
 * chunk_rows (a multiple of 32) / num_streams (1..16) of 0 take the autotuning table's
 * values for the layout (q1_tuning_get_streaming). */
q1_status q1_run_streaming(q1_ctx* ctx, q1_layout layout, const q1_columns* host, uint64_t rows,
                           int32_t cutoff_day, q1_placement placement, const q1_launch* launch,
                           uint64_t chunk_rows, int num_streams, q1_result* out);

/* --- autotuning table (new; replaces scripts/try_grid_params.py's optimum table) ----
 * q1_init reads tuning/gfx950.txt (or $Q1_TUNING_FILE): one shape per layout and
 * placement ("launch" lines) and, optionally, shapes tuned at several row counts
 * ("launch_at <rows> ..." lines).  A launch that leaves threads_per_block /
 * tuples_per_thread 0 takes the shape tuned at the row count nearest its own (by ratio),
 * halving tuples_per_thread for ranges too small to give 4 blocks per compute unit.
 * These are configuration calls: not to be made while other calls on the context, or a
 * server over it, are running. */
/* Sets the shape of every row count for (layout, placement). */
q1_status q1_tuning_set(q1_ctx* ctx, q1_layout layout, q1_placement placement,
                        const q1_launch* best);
/* The "launch" line's shape (the largest row count swept). */
q1_status q1_tuning_get(q1_ctx* ctx, q1_layout layout, q1_placement placement, q1_launch* out);
/* The shape a q1_run of `rows` rows with no caller shape takes (tables inside the
 * TPC-H value domain). */
q1_status q1_tuning_resolve(q1_ctx* ctx, q1_layout layout, q1_placement placement, uint64_t rows,
                            q1_launch* out);
q1_status q1_tuning_set_best_placement(q1_ctx* ctx, q1_layout layout, q1_placement placement);
/* The chunk rows and stream count q1_run_streaming takes for 0 ("streaming" lines of the
 * table; else 16M rows and 4 streams). */
q1_status q1_tuning_get_streaming(q1_ctx* ctx, q1_layout layout, uint64_t* chunk_rows, int* num_streams);

/* --- serving (new): continuous batching over HBM-resident tables ------------------ */
typedef struct q1_server q1_server;
typedef struct {
  uint32_t max_batch;          /* requests per batch launch (<= 4096) */
  uint32_t max_inflight;       /* batches on the device at once (streams) */
  uint32_t queue_capacity;     /* pending requests before Q1_EBUSY */
  uint32_t batch_window_us;    /* how long the scheduler may hold a batch back to fill it while
                                  another batch is on the device (0: launch at once) */
  uint64_t memory_budget;      /* admission: device bytes the in-flight work may pin -- the
                                  columns of every table with requests in flight (once per
                                  table) plus 480 B of descriptor and result per request;
                                  0 = three quarters of the device's memory.  A request whose
                                  table alone does not fit gets Q1_ENOMEM at submit. */
  uint64_t exclusive_bytes;    /* a request that scans at least this many column bytes is
                                  large: it runs alone (the scan kernel with its tuned launch),
                                  and large requests run one at a time, in arrival order, while
                                  smaller ones keep batching beside them; 0 = 512 MiB (about
                                  75 us of HBM time), UINT64_MAX = never */
} q1_server_config;

typedef struct {
  double submit_to_start_us;   /* queueing delay */
  double latency_us;           /* submit to result available */
  uint32_t batch_size;         /* requests in the batch that carried it */
  uint32_t batch_id;
} q1_request_stats;

q1_status q1_server_create(q1_ctx* ctx, const q1_server_config* cfg, q1_server** out);
q1_status q1_server_submit(q1_server* s, const q1_request* req, uint64_t* ticket);
q1_status q1_server_wait(q1_server* s, uint64_t ticket, int64_t timeout_us, q1_result* out,
                         q1_request_stats* stats);
q1_status q1_server_cancel(q1_server* s, uint64_t ticket);
q1_status q1_server_drain(q1_server* s);
void q1_server_destroy(q1_server* s);

#ifdef __cplusplus
}
#endif

#endif /* Q1_HIP_H_ */
