// The CPU/GPU split (src/main.cu:513-550 with the CoProc of src/cpu.h:23-38) and the
// host-resident streaming pipeline (src/main.cu:567-724): chunks of host columns are
// copied into per-stream device staging buffers and aggregated by the scan kernel on the
// same stream, so copies of one chunk overlap kernels of others; all launches add into
// one device table.
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

#include "q1_internal.h"

namespace q1 {

struct StreamState {
  q1_layout layout = Q1_LAYOUT_COMPRESSED;
  uint64_t chunk_rows = 0;
  int streams = 0;
  std::vector<hipStream_t> s;
  std::vector<hipEvent_t> done;
  std::vector<q1_columns> staging;
  std::vector<void*> allocs;
  std::vector<void*> workspace;
  uint64_t workspace_bytes = 0;
};

static void stream_state_free(StreamState* st) {
  if (!st) return;
  for (hipStream_t s : st->s) (void)hipStreamSynchronize(s);
  for (void* p : st->allocs) (void)hipFree(p);
  for (void* p : st->workspace)
    if (p) (void)hipFree(p);
  for (hipEvent_t e : st->done) (void)hipEventDestroy(e);
  for (hipStream_t s : st->s) (void)hipStreamDestroy(s);
  delete st;
}

void stream_state_release(q1_ctx* ctx) {
  stream_state_free(ctx->streaming);
  ctx->streaming = nullptr;
}

static void column_bytes_local(q1_layout layout, uint64_t n, uint64_t b[8]) {
  q1_layout_column_bytes(layout, n, b);
}

// The context's cached pipeline for (layout, chunk_rows, streams), rebuilt when those
// change.  A pipeline is stored only once all of its streams, events and staging buffers
// exist, so a failed build leaves nothing half-made behind.
static q1_status prepare(q1_ctx* ctx, q1_layout layout, uint64_t chunk_rows, int streams, StreamState** out) {
  StreamState* cur = ctx->streaming;
  if (cur && cur->layout == layout && cur->chunk_rows == chunk_rows && cur->streams == streams) {
    *out = cur;
    return Q1_OK;
  }
  stream_state_free(cur);
  ctx->streaming = nullptr;
  StreamState* st = new StreamState();
  auto give_up = [&](q1_status code, const char* why) {
    stream_state_free(st);
    return fail(ctx, code, why);
  };
  uint64_t b[8];
  column_bytes_local(layout, chunk_rows, b);
  for (int i = 0; i < streams; i++) {
    hipStream_t s;
    hipEvent_t e;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return give_up(Q1_EDEVICE, "stream creation failed");
    st->s.push_back(s);
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return give_up(Q1_EDEVICE, "event creation failed");
    st->done.push_back(e);
    q1_columns c{};
    const void** slots[8] = {&c.ship_date, &c.quantity, &c.extended_price, &c.discount,
                             &c.tax, &c.return_flag, &c.line_status, &c.filter};
    for (int k = 0; k < 8; k++) {
      if (!b[k]) continue;
      void* p = nullptr;
      if (hipMalloc(&p, b[k]) != hipSuccess) return give_up(Q1_ENOMEM, "staging allocation failed");
      st->allocs.push_back(p);
      *slots[k] = p;
    }
    st->staging.push_back(c);
    st->workspace.push_back(nullptr);
  }
  st->layout = layout;
  st->chunk_rows = chunk_rows;
  st->streams = streams;
  ctx->streaming = st;
  *out = st;
  return Q1_OK;
}

}  // namespace q1

using namespace q1;

extern "C" {

q1_status q1_run_streaming(q1_ctx* ctx, q1_layout layout, const q1_columns* host, uint64_t rows, int32_t cutoff_day,
                           q1_placement placement, const q1_launch* launch, uint64_t chunk_rows, int num_streams,
                           q1_result* out) {
  if (!ctx || !host || !out || layout < Q1_LAYOUT_REFERENCE || layout > Q1_LAYOUT_COMPRESSED_PUSHDOWN)
    return fail(ctx, Q1_EINVAL, "bad argument");
  // 0 = the autotuning table's value for the layout ("streaming" lines), else 16M rows / 4 streams
  if (chunk_rows == 0) chunk_rows = ctx->stream_chunk_rows[layout] ? ctx->stream_chunk_rows[layout] : kStreamChunkRows;
  if (num_streams == 0) num_streams = ctx->stream_count[layout] ? ctx->stream_count[layout] : kStreamCount;
  if (chunk_rows % 32 || num_streams < 1 || num_streams > 16)
    return fail(ctx, Q1_EINVAL, "chunk_rows must be a positive multiple of 32 and streams in [1, 16]");
  try {
    std::lock_guard<std::mutex> g(ctx->lock);
    const q1_placement p = resolve_placement(ctx, layout, placement);
    q1_launch l;
    resolve_launch(ctx, layout, p, chunk_rows, launch, &l, true);
    StreamState* st = nullptr;
    q1_status s = prepare(ctx, layout, chunk_rows, num_streams, &st);
    if (s != Q1_OK) return s;
    LaunchShape full_shape;
    const char* why = "";
    s = make_shape(p, layout, &l, 0, chunk_rows, &full_shape, &why);
    if (s != Q1_OK) return fail(ctx, s, why);
    const uint64_t ws_bytes = scan_workspace_bytes(p, full_shape);
    if (ws_bytes > st->workspace_bytes) {
      for (auto& w : st->workspace) {
        if (w) (void)hipFree(w);
        w = nullptr;
        if (hipMalloc(&w, ws_bytes) != hipSuccess) return fail(ctx, Q1_ENOMEM, "workspace allocation failed");
      }
      st->workspace_bytes = ws_bytes;
    }
    const bool fast = (l.flags & Q1_LAUNCH_ASSUME_TPCH_DOMAIN) != 0;
    uint64_t elem[8];
    column_bytes_local(layout, 1, elem);  // per-row bytes of the unpacked columns
    const bool packed = layout == Q1_LAYOUT_COMPRESSED || layout == Q1_LAYOUT_COMPRESSED_PUSHDOWN;
    const auto w0 = std::chrono::steady_clock::now();
    // Once copies are queued, an error return first drains every stream, so no copy is
    // still reading the caller's host columns when the call returns.
    auto drain_fail = [&](q1_status code, const char* why) {
      for (hipStream_t ss : st->s) (void)hipStreamSynchronize(ss);
      (void)hipStreamSynchronize(ctx->stream);
      return fail(ctx, code, why);
    };
    s = launch_zero(ctx->d_aggr, 1, ctx->stream);
    if (s != Q1_OK) return fail(ctx, s, "zero kernel failed");
    if (hipEventRecord(ctx->ev_begin, ctx->stream) != hipSuccess) return fail(ctx, Q1_EDEVICE, "event record failed");
    for (hipStream_t ss : st->s)
      if (hipStreamWaitEvent(ss, ctx->ev_begin, 0) != hipSuccess) return drain_fail(Q1_EDEVICE, "stream wait failed");
    const void* hsrc[8] = {host->ship_date, host->quantity, host->extended_price, host->discount,
                           host->tax, host->return_flag, host->line_status, host->filter};
    uint64_t bytes = 0;
    int idx = 0;
    for (uint64_t off = 0; off < rows; off += chunk_rows, idx = (idx + 1) % num_streams) {
      const uint64_t n = rows - off < chunk_rows ? rows - off : chunk_rows;
      const q1_columns& dst = st->staging[(size_t)idx];
      const void* ddst[8] = {dst.ship_date, dst.quantity, dst.extended_price, dst.discount,
                             dst.tax, dst.return_flag, dst.line_status, dst.filter};
      hipStream_t ss = st->s[(size_t)idx];
      for (int k = 0; k < 8; k++) {
        if (!ddst[k]) continue;
        uint64_t src_off, len;
        if (packed && k == 5) {          // return flags: 16 rows per word
          src_off = off / 16 * 4;
          len = (n + 15) / 16 * 4;
        } else if (packed && k >= 6) {   // line status / filter: 32 rows per word
          src_off = off / 32 * 4;
          len = (n + 31) / 32 * 4;
        } else {
          src_off = off * elem[k];
          len = n * elem[k];
        }
        if (!hsrc[k]) return drain_fail(Q1_EINVAL, "a column required by the layout is NULL");
        if (hipMemcpyAsync((void*)ddst[k], (const char*)hsrc[k] + src_off, len, hipMemcpyHostToDevice, ss) != hipSuccess)
          return drain_fail(Q1_EDEVICE, "chunk copy failed");
        bytes += len;
      }
      LaunchShape shape;
      s = make_shape(p, layout, &l, 0, n, &shape, &why);
      if (s != Q1_OK) return drain_fail(s, why);
      ScanArgs a;
      a.cols = dst;
      a.row_begin = 0;
      a.row_end = n;
      a.cutoff_day = cutoff_day;
      a.out = ctx->d_aggr;
      s = launch_scan(p, layout, fast, a, shape, st->workspace[(size_t)idx], ss);
      if (s != Q1_OK) return drain_fail(s, "scan kernel launch failed");
    }
    for (int i = 0; i < num_streams; i++) {
      if (hipEventRecord(st->done[(size_t)i], st->s[(size_t)i]) != hipSuccess ||
          hipStreamWaitEvent(ctx->stream, st->done[(size_t)i], 0) != hipSuccess)
        return drain_fail(Q1_EDEVICE, "stream join failed");
    }
    if (hipEventRecord(ctx->ev_end, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(ctx->h_aggr, ctx->d_aggr, sizeof(q1_device_aggregates), hipMemcpyDeviceToHost, ctx->stream) !=
            hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
      return drain_fail(Q1_EDEVICE, "final copy failed");
    q1_aggregates_to_result(ctx->h_aggr, out);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end);
    ctx->last = q1_timing{};
    ctx->last.kernel_ms = ms;
    ctx->last.h2d_ms = ms;
    ctx->last.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    ctx->last.rows = rows;
    ctx->last.bytes_read = bytes;
    ctx->last.threads_per_block = l.threads_per_block;
    ctx->last.tuples_per_thread = l.tuples_per_thread;
    ctx->last.placement = fast ? (uint32_t)p : (uint32_t)Q1_PLACE_IN_REGISTERS_PER_THREAD;
    return Q1_OK;
  } catch (const std::exception& e) {
    return fail(ctx, Q1_ENOMEM, e.what());
  }
}

q1_status q1_run_split(q1_ctx* ctx, const q1_table* t, const q1_columns* host, uint64_t rows, int32_t cutoff_day,
                       q1_placement placement, const q1_launch* launch, double cpu_fraction, uint64_t align_rows,
                       int cpu_threads, q1_result* out, double* cpu_ms, double* gpu_ms) {
  if (!ctx || !t || !host || !out || cpu_fraction < 0.0 || cpu_fraction > 1.0 || rows > t->rows)
    return fail(ctx, Q1_EINVAL, "bad argument");
  try {
    // split = rows - rows*fraction, rounded down to the alignment (src/main.cu:518-529);
    // packed columns additionally need a 32-row boundary.
    uint64_t split = rows - (uint64_t)((double)rows * cpu_fraction);
    uint64_t align = align_rows ? align_rows : 1;
    if (t->layout == Q1_LAYOUT_COMPRESSED || t->layout == Q1_LAYOUT_COMPRESSED_PUSHDOWN)
      align = (align + 31) / 32 * 32;
    split -= split % align;
    std::lock_guard<std::mutex> g(ctx->lock);
    const q1_placement p = resolve_placement(ctx, t->layout, placement);
    q1_launch l;
    resolve_launch(ctx, t->layout, p, split, launch, &l);
    LaunchShape shape;
    const char* why = "";
    q1_status s = make_shape(p, t->layout, &l, 0, split, &shape, &why);
    if (s != Q1_OK) return fail(ctx, s, why);
    s = ensure_workspace(ctx, scan_workspace_bytes(p, shape));
    if (s != Q1_OK) return s;
    ScanArgs a;
    a.cols = t->dev;
    a.row_begin = 0;
    a.row_end = split;
    a.cutoff_day = t->layout == Q1_LAYOUT_COMPRESSED_PUSHDOWN ? t->filter_cutoff : cutoff_day;
    a.out = ctx->d_aggr;
    const auto w0 = std::chrono::steady_clock::now();
    s = launch_zero(ctx->d_aggr, 1, ctx->stream);
    if (s == Q1_OK && hipEventRecord(ctx->ev_begin, ctx->stream) != hipSuccess) s = Q1_EDEVICE;
    if (s == Q1_OK) s = scan_on(ctx, p, t->layout, t->fast, a, shape, ctx->stream, !t->fast);
    if (s == Q1_OK && hipEventRecord(ctx->ev_end, ctx->stream) != hipSuccess) s = Q1_EDEVICE;
    if (s == Q1_OK &&
        hipMemcpyAsync(ctx->h_aggr, ctx->d_aggr, sizeof(q1_device_aggregates), hipMemcpyDeviceToHost, ctx->stream) !=
            hipSuccess)
      s = Q1_EDEVICE;
    if (s != Q1_OK) return fail(ctx, s, "device part of the split failed to launch");
    // The host share runs while the device works.
    q1_result cpu{};
    const auto c0 = std::chrono::steady_clock::now();
    s = q1_cpu_run(t->layout, host, split, rows - split, a.cutoff_day, cpu_threads, &cpu);
    const double cms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - c0).count();
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(ctx, Q1_EDEVICE, "device part of the split failed");
    if (s != Q1_OK) return fail(ctx, s, "host part of the split failed");
    q1_result gpu;
    q1_aggregates_to_result(ctx->h_aggr, &gpu);
    q1_result_merge(&gpu, &cpu, out);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end);
    ctx->last = q1_timing{};
    ctx->last.kernel_ms = ms;
    ctx->last.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    ctx->last.rows = rows;
    ctx->last.bytes_read = algorithmic_bytes(t->layout, split);
    ctx->last.num_blocks = shape.blocks;
    ctx->last.threads_per_block = shape.threads;
    ctx->last.tuples_per_thread = l.tuples_per_thread;
    ctx->last.placement = (uint32_t)p;
    if (cpu_ms) *cpu_ms = cms;
    if (gpu_ms) *gpu_ms = ms;
    return Q1_OK;
  } catch (const std::exception& e) {
    return fail(ctx, Q1_ENOMEM, e.what());
  }
}

}  // extern "C"
