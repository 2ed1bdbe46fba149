// Internal interface between the C-ABI layer (q1_api.cpp) and the gfx950 kernels
// (q1_kernels.hip).  Host-only declarations; no HIP types leak into the public header.
#pragma once

#include <stdint.h>

#include "../../include/q1_hip.h"

namespace q1 {

// Value domain of a table, measured on the device when the table is created.  Inside the
// TPC-H domain the narrow 32-bit kernels are exact; outside it the 64-bit kernels run.
struct Domain {
  int64_t min_disc, max_disc, min_tax, max_tax, min_qty, max_qty, min_price, max_price;
};

bool domain_is_tpch(q1_layout layout, const Domain& d);

// The blocks of a launch over rows [row_begin, row_end) start at row_begin rounded down to
// a multiple of kRowAlign rows (rows before row_begin are masked off in the first block).
// 128 rows are a whole number of tiles of every layout and kernel, and of 128-byte lines of
// the 1-, 2- and 4-byte columns, so every wavefront load of a column starts on a line,
// wherever the range starts: a range starting at an arbitrary row otherwise made every
// load touch one line more and ran 10 % slower.
constexpr uint64_t kRowAlign = 128;
inline uint64_t align_rows_down(uint64_t row) { return row / kRowAlign * kRowAlign; }

struct LaunchShape {
  uint32_t blocks, threads, tiles_per_thread;
};

// Kernel-level request: accumulate rows [row_begin, row_end) of `cols` into `out`.
struct ScanArgs {
  q1_columns cols;          // device pointers
  uint64_t row_begin, row_end;
  int32_t cutoff_day;       // MonetDB day number (ignored for pushdown)
  q1_device_aggregates* out;
  uint32_t out_mask;        // set by launch_scan: block b merges into out[b & out_mask]
};

// Spread merge.  Thousands of blocks adding into one 48-word table serialise on those
// words' atomics at the end of the grid; instead block b of a large launch merges into
// tables[b % kSpreadTables] and a one-block fold kernel adds the tables into the
// destination and clears them.  The tables must be zero before the launch (the fold
// leaves them zero) and must not be shared by launches that can overlap: one set per
// stream that serialises its launches.
constexpr uint32_t kSpreadTables = 64;      // power of two
// Smaller grids merge directly.  Device-scope atomics on one address from all compute
// units serialise: a grid's direct merge costs about 0.03 us per block against the fold's
// fixed 4.5 us, so direct merging wins below about 165-190 blocks (measured on 5K-6M-row
// launches: 49 blocks 11 us direct / 15 us spread, 489 blocks 28 / 15 us, 977 blocks 49 /
// 16 us).
constexpr uint32_t kSpreadMinBlocks = 176;
struct Spread {
  q1_device_aggregates* tables;  // kSpreadTables zeroed tables
};

// Validate and complete a launch shape for `rows` rows (0 fields filled from defaults).
q1_status make_shape(q1_placement placement, q1_layout layout, const q1_launch* launch,
                     uint64_t row_begin, uint64_t row_end, LaunchShape* shape, const char** why);

// Enqueue the scan kernel.  `fast` selects the narrow (TPC-H domain) kernels.  With a
// `spread` (may be null) a fast launch of at least kSpreadMinBlocks blocks merges through
// the spread tables and is followed by the fold kernel on the same stream.
// fast: the rows lie inside the TPC-H domain.  Otherwise an any-value kernel runs;
// table_outside picks the one for tables known to hold values outside the domain.
q1_status launch_scan(q1_placement placement, q1_layout layout, bool fast, const ScanArgs& a,
                      const LaunchShape& shape, void* workspace, void* stream, const Spread* spread = nullptr,
                      bool table_outside = false);
uint64_t scan_workspace_bytes(q1_placement placement, const LaunchShape& shape);

// Zero a device aggregate table (a fill kernel, safe inside graph capture).
q1_status launch_zero(q1_device_aggregates* table, uint32_t count, void* stream);

// Device-side generator: rows [0, rows) of `layout` into the device columns.
q1_status launch_generate(q1_layout layout, uint64_t seed, uint64_t rows, const q1_columns& cols,
                          void* stream);

// Device-side domain statistics of a table into `out_dev` (8 int64 slots: min/max pairs).
q1_status launch_domain(q1_layout layout, const q1_columns& cols, uint64_t rows, int64_t* out_dev,
                        void* stream);

// Device-side pushdown producer: filter bitmap of a compressed ship_date column.
q1_status launch_build_filter(const uint16_t* ship, uint64_t rows, int32_t cutoff_day,
                              uint32_t* filter, void* stream);

// Batched scan: one launch over many requests of one layout, TPC-H domain.  The device
// buffer starts with a BatchHeader followed by `n` BatchDesc records.  Request r is cut
// into the blocks launch_scan would run over its rows alone (256 lanes x
// tiles_per_thread tiles each); the blocks of all requests, in request order, form one
// list of total_blocks entries.  A resident grid of G blocks works through the list with a
// fixed stride (block b runs entries b, b + G, b + 2G, ...).  Entry v of request r merges
// into spread table v mod 64 of the request (spread + 64 r) when r < spread_requests, and
// into outs[r] otherwise; after the kernel, fold_spread (one block per request with spread
// tables) adds each request's 64 tables into outs[r] and clears them.
struct BatchDesc {
  q1_columns cols;
  uint64_t row_begin, row_end;
  uint64_t block_begin;       // request's first entry in the block list
  int32_t cutoff_day;
  uint32_t pad;
};
struct BatchHeader {
  uint64_t n;                 // requests
  uint64_t total_blocks;      // entries of the block list (below 2^31)
  uint32_t tiles_per_thread;  // tiles of layout_tile_rows rows per lane and entry
  uint32_t spread_requests;   // requests 0 .. spread_requests-1 merge through spread tables
  uint64_t pad[5];
};
constexpr uint32_t kBatchThreads = 256;
// Requests of a batch that get spread tables: 64 x 384 B each, at most 24 MiB.
constexpr uint32_t kBatchSpreadRequests = 1024;
uint32_t layout_tile_rows(q1_layout layout);
// The placement the batched kernel runs for a requested one (shared or in registers),
// the largest rows per lane it accepts for it, and how many of its blocks fit on a CU.
q1_placement batch_placement(q1_placement p);
uint32_t batch_max_lane_rows(q1_placement p);
uint32_t batch_blocks_per_cu(q1_layout layout, q1_placement p);
// The batched kernel over the list in dev_buffer, then, with `spread` (spread_requests x
// kSpreadTables tables, all zero), the fold of the first spread_requests requests' tables:
// added into outs, or, with store_results (a device-visible pointer to the results in
// pinned host memory, when every request has spread tables), stored there directly.
q1_status launch_batch(q1_layout layout, q1_placement placement, BatchHeader* dev_buffer,
                       q1_device_aggregates* outs, q1_device_aggregates* spread, uint32_t spread_requests,
                       q1_device_aggregates* store_results, uint32_t blocks, uint32_t threads, void* stream);
// The host function launch_batch launches for (layout, placement), e.g. to find its node in
// a captured graph.
const void* batch_kernel_function(q1_layout layout, q1_placement placement);

// The scratch of q1_enqueue_noalloc (q1_noalloc.hip): a static per-device array holding the
// spread tables and kNoallocWorkspaceBytes of per-block tables for the global placement.
// noalloc_setup looks up the array's address on `device` (which must be current) and
// clears its spread tables, once per device; noalloc_launch_scan is launch_scan with the
// array as its scratch, refusing launches that need more (*why says which).
constexpr uint64_t kNoallocWorkspaceBytes = 8ull << 20;
q1_status noalloc_setup(int device, void* stream);
q1_status noalloc_launch_scan(int device, q1_placement placement, q1_layout layout, bool fast, const ScanArgs& a,
                              const LaunchShape& shape, void* stream, const char** why);

}  // namespace q1
