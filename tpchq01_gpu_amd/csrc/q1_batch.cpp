// q1_run_batch: many filter-aggregate requests (uneven row counts, their own cutoffs) in
// one persistent launch, captured once into a hipGraph and replayed:
//   H2D copy of the request descriptors (pinned staging) -> the batched kernel -> the fold
//   of each request's spread tables, which stores the results straight into the pinned
//   host results (over PCIe).
// A batch of more than kBatchSpreadRequests requests gives spread tables to the first
// kBatchSpreadRequests only, so its graph is: H2D copy -> zero the device results (a fill
// kernel, never a memset node) -> the batched kernel -> the fold, adding into the device
// results -> D2H copy of the results.
// The spread tables are zeroed once when allocated; every fold clears the tables it read.
// The graph is rebuilt only when the capacity bucket, layout, placement or grid changes;
// each call only rewrites the pinned descriptor buffer before the replay.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "q1_internal.h"

struct q1_batch_state {
  q1_layout layout = Q1_LAYOUT_COMPRESSED;
  q1_placement placement = Q1_PLACE_IN_REGISTERS;
  uint32_t capacity = 0, blocks = 0, threads = 0;
  q1::BatchHeader* h_buf = nullptr;  // pinned: header + capacity descriptors
  q1::BatchHeader* d_buf = nullptr;
  q1_device_aggregates* h_out = nullptr;  // pinned
  q1_device_aggregates* d_out = nullptr;
  q1_device_aggregates* d_spread = nullptr;  // spread_requests x kSpreadTables tables
  uint32_t spread_requests = 0;
  q1_device_aggregates* h_out_dev = nullptr;  // h_out as the device addresses it (folds store there)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace q1 {

void batch_destroy(q1_batch_state* b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->h_buf) (void)hipHostFree(b->h_buf);
  if (b->h_out) (void)hipHostFree(b->h_out);
  if (b->d_buf) (void)hipFree(b->d_buf);
  if (b->d_out) (void)hipFree(b->d_out);
  if (b->d_spread) (void)hipFree(b->d_spread);
  delete b;
}

static uint64_t buf_bytes(uint32_t cap) { return sizeof(BatchHeader) + (uint64_t)cap * sizeof(BatchDesc); }

// Rows per lane: the autotuner's batch value for the layout and placement, else the
// scan's tuned value, within the batched kernel's limit.
uint32_t batch_lane_rows(q1_ctx* ctx, q1_layout layout, q1_placement p) {
  p = batch_placement(p);
  const uint32_t most = batch_max_lane_rows(p);
  const uint32_t batch_tuned = ctx->batch_tuning[layout][p].tuples_per_thread;
  if (batch_tuned >= 16 && batch_tuned <= most) return batch_tuned / 16 * 16;
  const uint32_t tuned = ctx->tuning[layout][p].tuples_per_thread;
  return tuned >= 16 && tuned <= most ? tuned / 16 * 16 : std::min<uint32_t>(256, most);
}

// Grid: the autotuner's blocks per compute unit for the layout and placement, else as many
// blocks as are resident at once.
uint32_t batch_default_blocks(q1_ctx* ctx, q1_layout layout, q1_placement p) {
  p = batch_placement(p);
  const uint32_t cus = (uint32_t)(ctx->compute_units > 0 ? ctx->compute_units : 256);
  const uint32_t per_cu = ctx->batch_tuning[layout][p].num_blocks;
  return cus * (per_cu ? per_cu : batch_blocks_per_cu(layout, p));
}

uint64_t batch_describe(BatchDesc* d, const q1_request& r, uint64_t first, uint32_t lane_rows) {
  const q1_table* t = r.table;
  d->cols = t->dev;
  d->row_begin = r.row_begin;
  d->row_end = r.row_begin + r.row_count;
  d->block_begin = first;
  d->cutoff_day = t->layout == Q1_LAYOUT_COMPRESSED_PUSHDOWN ? t->filter_cutoff : r.cutoff_day;
  d->pad = 0;
  if (r.row_count == 0) return 0;
  const uint64_t trows = layout_tile_rows(t->layout);
  const uint64_t tiles = (d->row_end + trows - 1) / trows - align_rows_down(d->row_begin) / trows;
  const uint64_t per_block = (uint64_t)kBatchThreads * (lane_rows / trows);
  return (tiles + per_block - 1) / per_block;
}

static q1_status build_state(q1_ctx* ctx, q1_layout layout, q1_placement placement, uint32_t cap,
                             uint32_t threads, uint32_t blocks) {
  q1_batch_state* b = ctx->batch;
  if (b && b->layout == layout && b->placement == placement && b->capacity == cap && b->threads == threads &&
      b->blocks == blocks)
    return Q1_OK;
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(ctx, Q1_EDEVICE, "stream sync failed");
  batch_destroy(b);
  ctx->batch = nullptr;
  b = new q1_batch_state();
  b->layout = layout;
  b->placement = placement;
  b->capacity = cap;
  b->threads = threads;
  b->blocks = blocks;
  b->spread_requests = std::min(cap, kBatchSpreadRequests);
  const uint64_t spread_bytes = (uint64_t)b->spread_requests * kSpreadTables * sizeof(q1_device_aggregates);
  bool ok = hipHostMalloc(&b->h_buf, buf_bytes(cap), hipHostMallocDefault) == hipSuccess &&
            hipHostMalloc(&b->h_out, cap * sizeof(q1_device_aggregates), hipHostMallocMapped) == hipSuccess &&
            hipHostGetDevicePointer((void**)&b->h_out_dev, b->h_out, 0) == hipSuccess &&
            hipMalloc(&b->d_buf, buf_bytes(cap)) == hipSuccess &&
            hipMalloc(&b->d_out, cap * sizeof(q1_device_aggregates)) == hipSuccess &&
            hipMalloc(&b->d_spread, spread_bytes) == hipSuccess &&
            hipMemsetAsync(b->d_spread, 0, spread_bytes, ctx->stream) == hipSuccess &&
            hipStreamSynchronize(ctx->stream) == hipSuccess;  // before the capture, not inside it
  if (!ok) {
    batch_destroy(b);
    return fail(ctx, Q1_ENOMEM, "batch buffers could not be allocated");
  }
  std::memset(b->h_buf, 0, buf_bytes(cap));
  // Capture the whole step once.
  hipStream_t s = ctx->stream;
  q1_status st = Q1_OK;
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) st = Q1_EDEVICE;
  if (st == Q1_OK && hipMemcpyAsync(b->d_buf, b->h_buf, buf_bytes(cap), hipMemcpyHostToDevice, s) != hipSuccess)
    st = Q1_EDEVICE;
  const bool store = b->spread_requests == cap;  // every request has spread tables
  if (st == Q1_OK && !store) st = launch_zero(b->d_out, cap, s);
  if (st == Q1_OK)
    st = launch_batch(layout, placement, b->d_buf, b->d_out, b->d_spread, b->spread_requests,
                      store ? b->h_out_dev : nullptr, blocks, threads, s);
  if (st == Q1_OK && !store &&
      hipMemcpyAsync(b->h_out, b->d_out, cap * sizeof(q1_device_aggregates), hipMemcpyDeviceToHost, s) != hipSuccess)
    st = Q1_EDEVICE;
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(s, &g);
  if (st != Q1_OK || ec != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    batch_destroy(b);
    return fail(ctx, Q1_EDEVICE, "capturing the batch graph failed");
  }
  b->graph = g;
  if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
    batch_destroy(b);
    return fail(ctx, Q1_EDEVICE, "instantiating the batch graph failed");
  }
  ctx->batch = b;
  return Q1_OK;
}

}  // namespace q1

using namespace q1;

extern "C" q1_status q1_run_batch(q1_ctx* ctx, const q1_request* reqs, uint32_t n, const q1_launch* launch,
                                  q1_result* outs) {
  if (!ctx || (!reqs && n) || (!outs && n)) return fail(ctx, Q1_EINVAL, "bad argument");
  if (n == 0) return Q1_OK;
  if (n > (1u << 16)) return fail(ctx, Q1_EINVAL, "at most 65536 requests per batch");
  for (uint32_t i = 0; i < n; i++) {
    const q1_table* t = reqs[i].table;
    if (!t) return fail(ctx, Q1_EINVAL, "request without a table");
    if (reqs[i].row_begin > t->rows || reqs[i].row_count > t->rows - reqs[i].row_begin)
      return fail(ctx, Q1_EINVAL, "request row range outside its table");
  }
  try {
    // One persistent launch needs one layout and the TPC-H domain; any other mix runs
    // request by request (same results, no batching).
    const q1_layout layout = reqs[0].table->layout;
    bool batchable = true;
    for (uint32_t i = 0; i < n; i++)
      batchable = batchable && reqs[i].table->layout == layout && reqs[i].table->fast;
    if (!batchable) {
      for (uint32_t i = 0; i < n; i++) {
        const q1_status s = q1_run(ctx, reqs[i].table, reqs[i].row_begin, reqs[i].row_count, reqs[i].cutoff_day,
                                   Q1_PLACE_AUTO, nullptr, &outs[i]);
        if (s != Q1_OK) return s;
      }
      return Q1_OK;
    }
    std::lock_guard<std::mutex> g(ctx->lock);
    const q1_placement p = batch_placement(ctx->best[layout]);
    // The batched kernel runs 256-thread blocks; the grid and rows per lane are tunable.
    const uint32_t threads = kBatchThreads;
    if (launch && launch->threads_per_block && launch->threads_per_block != threads)
      return fail(ctx, Q1_EINVAL, "the batched kernel runs 256 threads per block");
    uint32_t lane_rows = batch_lane_rows(ctx, layout, p);
    if (launch && launch->tuples_per_thread) {
      lane_rows = launch->tuples_per_thread;
      if (lane_rows % 16 != 0 || lane_rows > batch_max_lane_rows(p))
        return fail(ctx, Q1_EINVAL,
                    "batched tuples_per_thread must be a multiple of 16, at most 256 (shared placement) or 16384");
    }
    const uint32_t blocks =
        launch && launch->num_blocks ? launch->num_blocks : batch_default_blocks(ctx, layout, p);
    // The grid's work-items (blocks x 256) must stay below 2^32.
    if (blocks >= (1u << 24)) return fail(ctx, Q1_EINVAL, "batched num_blocks must be below 2^24");
    uint32_t cap = 16;
    while (cap < n) cap <<= 1;
    q1_status s = build_state(ctx, layout, p, cap, threads, blocks);
    if (s != Q1_OK) return s;
    q1_batch_state* b = ctx->batch;
    BatchDesc* d = reinterpret_cast<BatchDesc*>(b->h_buf + 1);
    uint64_t entries = 0, rows = 0;
    for (uint32_t i = 0; i < n; i++) {
      entries += batch_describe(&d[i], reqs[i], entries, lane_rows);
      rows += reqs[i].row_count;
    }
    if (entries >= (1ull << 31)) return fail(ctx, Q1_EINVAL, "batch too large (2^31 blocks of rows)");
    b->h_buf->n = n;
    b->h_buf->total_blocks = entries;
    b->h_buf->tiles_per_thread = lane_rows / layout_tile_rows(layout);
    b->h_buf->spread_requests = b->spread_requests;
    const auto w0 = std::chrono::steady_clock::now();
    if (hipEventRecord(ctx->ev_begin, ctx->stream) != hipSuccess ||
        hipGraphLaunch(b->exec, ctx->stream) != hipSuccess || hipEventRecord(ctx->ev_end, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
      // The spread tables may hold partial sums: the next call rebuilds the state.
      (void)hipStreamSynchronize(ctx->stream);
      batch_destroy(ctx->batch);
      ctx->batch = nullptr;
      return fail(ctx, Q1_EDEVICE, "batch graph launch failed");
    }
    for (uint32_t i = 0; i < n; i++) q1_aggregates_to_result(&b->h_out[i], &outs[i]);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end);
    ctx->last = q1_timing{};
    ctx->last.kernel_ms = ms;
    ctx->last.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    ctx->last.rows = rows;
    ctx->last.bytes_read = algorithmic_bytes(layout, rows);
    ctx->last.num_blocks = blocks;
    ctx->last.threads_per_block = threads;
    ctx->last.tuples_per_thread = lane_rows;
    ctx->last.placement = (uint32_t)p;
    return Q1_OK;
  } catch (const std::exception& e) {
    return fail(ctx, Q1_ENOMEM, e.what());
  }
}
