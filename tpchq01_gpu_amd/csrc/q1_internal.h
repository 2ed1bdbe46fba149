// Internal host-side state shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/q1_hip.h"
#include "q1_kernels.h"

struct q1_table {
  q1_ctx* ctx = nullptr;
  q1_layout layout = Q1_LAYOUT_COMPRESSED;
  uint64_t rows = 0;
  q1_columns dev{};
  void* owned[8] = {};           // allocations this table frees
  q1::Domain domain{};
  bool fast = false;             // inside the TPC-H value domain
  uint64_t device_bytes = 0;
  int32_t filter_cutoff = 0;     // pushdown tables: the cutoff their bitmap encodes
};

// q1_run_streaming's chunk rows and stream count when neither the caller nor the tuning
// table gives them.
constexpr uint64_t kStreamChunkRows = 1ull << 24;
constexpr int kStreamCount = 4;

struct q1_batch_state;  // q1_batch.cpp
namespace q1 {
struct StreamState;     // q1_stream.cpp: the host-resident streaming pipeline
// Spread tables of one caller stream (q1_enqueue on a stream other than the context's).
// `done` is recorded after every launch that uses them and the next launch waits for it,
// so a stream created at the address of a destroyed one never overlaps the old stream's
// last use; `usable` is cleared if that bookkeeping ever fails.
struct StreamSpread {
  Spread spread{};
  hipEvent_t done = nullptr;
  bool dirty = true;   // zeroed by the first launch that uses them
  bool usable = true;
};
// Caller streams that get spread tables; launches on further streams merge directly.
constexpr size_t kMaxStreamSpreads = 16;
// A launch shape tuned on ranges of `rows` rows ("launch_at" lines of the tuning file).
struct SizedLaunch {
  uint64_t rows;
  uint32_t threads_per_block, tuples_per_thread;
};
}

struct q1_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_begin = nullptr, ev_end = nullptr, ev_h2d_begin = nullptr, ev_h2d_end = nullptr;
  q1_device_aggregates* d_aggr = nullptr;  // one table for q1_run
  q1_device_aggregates* h_aggr = nullptr;  // pinned host copy
  int64_t* d_stats = nullptr;              // 8 words for domain statistics
  void* workspace = nullptr;
  uint64_t workspace_bytes = 0;
  q1_timing last{};
  std::string error;
  int32_t compute_units = 0;
  q1_launch tuning[4][5]{};                // [layout][placement]
  // Shapes tuned at several row counts, ascending; a launch takes the one tuned nearest its
  // own row count (by ratio) and falls back to tuning[][] when there is none.
  std::vector<q1::SizedLaunch> tuning_at[4][5];
  // The batched kernel's tuned shape per [layout][placement]: num_blocks holds blocks per
  // compute unit, tuples_per_thread rows per lane; zero = not tuned.
  q1_launch batch_tuning[4][5]{};
  // The host-resident streaming pipeline's tuned chunk rows and stream count per layout
  // (0 = not tuned: kStreamChunkRows, kStreamCount).
  uint64_t stream_chunk_rows[4] = {};
  int stream_count[4] = {};
  q1_placement best[4] = {Q1_PLACE_IN_REGISTERS, Q1_PLACE_IN_REGISTERS, Q1_PLACE_IN_REGISTERS,
                          Q1_PLACE_IN_REGISTERS};
  q1_batch_state* batch = nullptr;
  q1::StreamState* streaming = nullptr;    // cached staging buffers and streams of q1_run_streaming
  q1::Spread spread{};                     // spread-merge tables of the context stream
  bool spread_dirty = false;               // a failed launch may have left partials in them
  std::unordered_map<void*, q1::StreamSpread> stream_spreads;  // caller streams' spread tables
  std::unordered_set<q1_table*> tables;    // live tables (released device memory at shutdown)
  std::mutex lock;                         // serialises calls that share the context stream
};

namespace q1 {
q1_status fail(q1_ctx* ctx, q1_status s, const std::string& why);
q1_status hip_fail(q1_ctx* ctx, hipError_t e, const char* what);
q1_placement resolve_placement(q1_ctx* ctx, q1_layout layout, q1_placement p);
// Fills a launch from the caller's, with the tuned values where it leaves them 0.  Launches
// without Q1_LAUNCH_ASSUME_TPCH_DOMAIN that run the any-value kernel (q1_enqueue,
// q1_run_streaming) take that kernel's own defaults instead: it runs every placement
// alike, and is fastest with 256 threads and 1024 rows per lane.
void resolve_launch(q1_ctx* ctx, q1_layout layout, q1_placement p, uint64_t rows, const q1_launch* in,
                    q1_launch* out, bool any_value_kernel = false);
uint64_t algorithmic_bytes(q1_layout layout, uint64_t rows);
q1_status ensure_workspace(q1_ctx* ctx, uint64_t bytes);
// launch_scan with the context's spread tables when `stream` is the context stream.
// table_outside: the rows come from a table known to hold values outside the domain.
q1_status scan_on(q1_ctx* ctx, q1_placement placement, q1_layout layout, bool fast, const ScanArgs& a,
                  const LaunchShape& shape, void* stream, bool table_outside = false);
void batch_destroy(q1_batch_state* b);
// The batched kernel's plan, shared by q1_run_batch and the server: rows per lane (the
// tuned value of the layout and placement, at most batch_max_lane_rows), the default
// grid (every block resident at once), and descriptor r of a batch, whose blocks start at
// list entry `first`; batch_describe returns the request's number of blocks.
uint32_t batch_lane_rows(q1_ctx* ctx, q1_layout layout, q1_placement p);
uint32_t batch_default_blocks(q1_ctx* ctx, q1_layout layout, q1_placement p);
uint64_t batch_describe(BatchDesc* d, const q1_request& r, uint64_t first, uint32_t lane_rows);
void stream_state_release(q1_ctx* ctx);
}  // namespace q1
