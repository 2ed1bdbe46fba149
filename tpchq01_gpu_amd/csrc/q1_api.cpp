// C-ABI implementation: contexts, device-resident tables, the query entry points,
// timing, autotuning table, the CPU/GPU split and the host-resident streaming pipeline.
// Every entry point catches C++ exceptions and returns a status (nothing crosses the ABI).
#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <thread>
#include <vector>

#include "q1_common.h"
#include "q1_internal.h"

namespace q1 {

q1_status fail(q1_ctx* ctx, q1_status s, const std::string& why) {
  if (ctx) ctx->error = why;
  return s;
}

q1_status hip_fail(q1_ctx* ctx, hipError_t e, const char* what) {
  if (ctx) ctx->error = std::string(what) + ": " + hipGetErrorString(e);
  return Q1_EDEVICE;
}

#define Q1_HIP(ctx, call)                                   \
  do {                                                      \
    const hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) return q1::hip_fail(ctx, e_, #call); \
  } while (0)

static bool valid_layout(int l) { return l >= Q1_LAYOUT_REFERENCE && l <= Q1_LAYOUT_COMPRESSED_PUSHDOWN; }
static bool valid_placement(int p) { return p >= Q1_PLACE_IN_REGISTERS && p <= Q1_PLACE_AUTO; }

uint64_t algorithmic_bytes(q1_layout layout, uint64_t rows) {
  // bits per row: 304 reference, 160 uncompressed, 75 compressed, 60 pushdown
  static const uint64_t bits[4] = {304, 160, 75, 60};
  return rows * bits[layout] / 8;
}

q1_placement resolve_placement(q1_ctx* ctx, q1_layout layout, q1_placement p) {
  return p == Q1_PLACE_AUTO ? ctx->best[layout] : p;
}

// The tuned shape for a range of `rows` rows: of the shapes tuned at several row counts,
// the one whose row count is nearest by ratio (the larger on a tie); the single tuned
// shape when the table has none.
static q1_launch tuned_for_rows(const q1_ctx* ctx, q1_layout layout, q1_placement p, uint64_t rows) {
  q1_launch t = ctx->tuning[layout][p];
  const std::vector<SizedLaunch>& at = ctx->tuning_at[layout][p];
  if (at.empty()) return t;
  const double r = (double)std::max<uint64_t>(rows, 1);
  double best = 0;
  for (size_t i = 0; i < at.size(); i++) {
    const double d = std::fabs(std::log(r / (double)at[i].rows));
    if (i == 0 || d <= best) {
      best = d;
      t.threads_per_block = at[i].threads_per_block;
      t.tuples_per_thread = at[i].tuples_per_thread;
    }
  }
  return t;
}

void resolve_launch(q1_ctx* ctx, q1_layout layout, q1_placement p, uint64_t rows, const q1_launch* in,
                    q1_launch* out, bool any_value_kernel) {
  const q1_launch t = tuned_for_rows(ctx, layout, p, rows);
  const bool any_value = any_value_kernel && !(in && (in->flags & Q1_LAUNCH_ASSUME_TPCH_DOMAIN));
  const uint32_t tuned_threads = any_value ? 256 : t.threads_per_block;
  const uint32_t tuned_tuples = any_value ? 1024 : t.tuples_per_thread;
  out->threads_per_block = in && in->threads_per_block ? in->threads_per_block : tuned_threads;
  out->tuples_per_thread = in && in->tuples_per_thread ? in->tuples_per_thread : tuned_tuples;
  out->num_blocks = in ? in->num_blocks : 0;
  out->flags = in ? in->flags : 0;
  // A range smaller than the row counts the shape was tuned at halves its tuples_per_thread
  // until the grid has at least 4 blocks per compute unit, so small requests still spread
  // over the whole chip.  A caller-chosen value is kept as given.
  if (!(in && in->tuples_per_thread) && !out->num_blocks) {
    const uint64_t want = 4ull * (uint64_t)(ctx->compute_units > 0 ? ctx->compute_units : 256);
    const uint64_t tpb = out->threads_per_block ? out->threads_per_block : 256;
    while (out->tuples_per_thread > 16 && rows < want * tpb * out->tuples_per_thread)
      out->tuples_per_thread = std::max<uint32_t>(16, (out->tuples_per_thread / 2) & ~15u);
  }
}

q1_status ensure_workspace(q1_ctx* ctx, uint64_t bytes) {
  if (bytes <= ctx->workspace_bytes) return Q1_OK;
  if (ctx->workspace) {
    Q1_HIP(ctx, hipStreamSynchronize(ctx->stream));
    Q1_HIP(ctx, hipFree(ctx->workspace));
    ctx->workspace = nullptr;
    ctx->workspace_bytes = 0;
  }
  Q1_HIP(ctx, hipMalloc(&ctx->workspace, bytes));
  ctx->workspace_bytes = bytes;
  return Q1_OK;
}

// The spread tables of a caller's stream for a large fast launch, made on the stream's
// first such launch; nullptr (merge directly) while the stream is being captured into a
// graph, where nothing may be allocated or waited for outside the capture, once
// kMaxStreamSpreads streams have tables, or if making them fails.  A stream that has
// tables first waits for the event of their last use.
static StreamSpread* stream_spread(q1_ctx* ctx, void* stream) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing((hipStream_t)stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return nullptr;
  auto it = ctx->stream_spreads.find(stream);
  if (it != ctx->stream_spreads.end()) {
    StreamSpread& ss = it->second;
    if (!ss.usable) return nullptr;
    if (hipStreamWaitEvent((hipStream_t)stream, ss.done, 0) != hipSuccess) {
      ss.usable = false;
      return nullptr;
    }
    return &ss;
  }
  if (ctx->stream_spreads.size() >= kMaxStreamSpreads) return nullptr;
  StreamSpread ss;
  if (hipMalloc(&ss.spread.tables, kSpreadTables * sizeof(q1_device_aggregates)) != hipSuccess) return nullptr;
  if (hipEventCreateWithFlags(&ss.done, hipEventDisableTiming) != hipSuccess) {
    (void)hipFree(ss.spread.tables);
    return nullptr;
  }
  return &ctx->stream_spreads.emplace(stream, ss).first->second;
}

q1_status scan_on(q1_ctx* ctx, q1_placement placement, q1_layout layout, bool fast, const ScanArgs& a,
                  const LaunchShape& shape, void* stream, bool table_outside) {
  // Launches on one stream run one after another, so they can share one set of spread
  // tables: the context's for the context stream, a set of the stream's own for a
  // caller's stream (large fast launches only; the others merge directly).
  Spread* sp = nullptr;
  bool* dirty = nullptr;
  StreamSpread* other = nullptr;
  if (stream == (void*)ctx->stream) {
    if (ctx->spread.tables) {
      sp = &ctx->spread;
      dirty = &ctx->spread_dirty;
    }
  } else if (fast && shape.blocks >= kSpreadMinBlocks) {
    if ((other = stream_spread(ctx, stream))) {
      sp = &other->spread;
      dirty = &other->dirty;
    }
  }
  if (sp && *dirty) {
    const q1_status z = launch_zero(sp->tables, kSpreadTables, stream);
    if (z != Q1_OK) return z;
    *dirty = false;
  }
  const q1_status s = launch_scan(placement, layout, fast, a, shape, ctx->workspace, stream, sp, table_outside);
  if (s != Q1_OK && sp) *dirty = true;
  if (other && hipEventRecord(other->done, (hipStream_t)stream) != hipSuccess) other->usable = false;
  return s;
}

// Column sizes in bytes, in q1_columns order.
static void column_bytes(q1_layout layout, uint64_t n, uint64_t b[8]) {
  std::memset(b, 0, 8 * sizeof(uint64_t));
  const uint64_t rf_words = (n + 15) / 16, ls_words = (n + 31) / 32;
  switch (layout) {
    case Q1_LAYOUT_REFERENCE:
      b[0] = 4 * n; b[1] = 8 * n; b[2] = 8 * n; b[3] = 8 * n; b[4] = 8 * n; b[5] = n; b[6] = n;
      break;
    case Q1_LAYOUT_UNCOMPRESSED:
      b[0] = 4 * n; b[1] = 4 * n; b[2] = 8 * n; b[3] = n; b[4] = n; b[5] = n; b[6] = n;
      break;
    case Q1_LAYOUT_COMPRESSED:
      b[0] = 2 * n; b[1] = n; b[2] = 4 * n; b[3] = n; b[4] = n; b[5] = 4 * rf_words; b[6] = 4 * ls_words;
      break;
    case Q1_LAYOUT_COMPRESSED_PUSHDOWN:
      b[1] = n; b[2] = 4 * n; b[3] = n; b[4] = n; b[5] = 4 * rf_words; b[6] = 4 * ls_words; b[7] = 4 * ls_words;
      break;
  }
}

static const void*& col(q1_columns& c, int i) {
  const void** p[8] = {&c.ship_date, &c.quantity, &c.extended_price, &c.discount,
                       &c.tax, &c.return_flag, &c.line_status, &c.filter};
  return *p[i];
}
static const void* colc(const q1_columns& c, int i) { return col(const_cast<q1_columns&>(c), i); }

static q1_status measure_domain(q1_ctx* ctx, q1_table* t) {
  const int64_t init[8] = {INT64_MAX, INT64_MIN, INT64_MAX, INT64_MIN,
                           INT64_MAX, INT64_MIN, INT64_MAX, INT64_MIN};
  int64_t out[8];
  Q1_HIP(ctx, hipMemcpyAsync(ctx->d_stats, init, sizeof(init), hipMemcpyHostToDevice, ctx->stream));
  q1_status s = launch_domain(t->layout, t->dev, t->rows, ctx->d_stats, ctx->stream);
  if (s != Q1_OK) return fail(ctx, s, "domain statistics kernel launch failed");
  Q1_HIP(ctx, hipMemcpyAsync(out, ctx->d_stats, sizeof(out), hipMemcpyDeviceToHost, ctx->stream));
  Q1_HIP(ctx, hipStreamSynchronize(ctx->stream));
  t->domain = Domain{out[0], out[1], out[2], out[3], out[4], out[5], out[6], out[7]};
  t->fast = t->rows == 0 || domain_is_tpch(t->layout, t->domain);
  return Q1_OK;
}

static void load_tuning_file(q1_ctx* ctx) {
  std::string path;
  if (const char* env = std::getenv("Q1_TUNING_FILE")) {
    path = env;
  } else {
    Dl_info info;
    if (dladdr((void*)&load_tuning_file, &info) && info.dli_fname) {
      std::string so = info.dli_fname;
      const size_t slash = so.rfind('/');
      const std::string dir = slash == std::string::npos ? "." : so.substr(0, slash);
      path = dir + "/../tuning/gfx950.txt";
    }
  }
  std::ifstream in(path);
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    std::string kind;
    ls >> kind;
    int l = -1, p = -1;
    if (kind == "launch") {
      uint32_t tpb = 0, tpt = 0;
      ls >> l >> p >> tpb >> tpt;
      // same limits as make_shape: a bad line is ignored rather than failing every launch
      const bool shape_ok = (tpb == 64 || tpb == 128 || tpb == 256) && tpt && tpt % 16 == 0 && tpt <= 16384;
      if (valid_layout(l) && p >= 0 && p <= Q1_PLACE_GLOBAL && shape_ok) {
        ctx->tuning[l][p].threads_per_block = tpb;
        ctx->tuning[l][p].tuples_per_thread = tpt;
      }
    } else if (kind == "launch_at") {
      // launch_at <rows> <layout> <placement> <threads_per_block> <tuples_per_thread>
      uint64_t rows = 0;
      uint32_t tpb = 0, tpt = 0;
      ls >> rows >> l >> p >> tpb >> tpt;
      const bool shape_ok = (tpb == 64 || tpb == 128 || tpb == 256) && tpt && tpt % 16 == 0 && tpt <= 16384;
      if (rows && valid_layout(l) && p >= 0 && p <= Q1_PLACE_GLOBAL && shape_ok) {
        std::vector<SizedLaunch>& at = ctx->tuning_at[l][p];
        auto it = std::lower_bound(at.begin(), at.end(), rows,
                                   [](const SizedLaunch& a, uint64_t r) { return a.rows < r; });
        if (it != at.end() && it->rows == rows)
          *it = SizedLaunch{rows, tpb, tpt};
        else
          at.insert(it, SizedLaunch{rows, tpb, tpt});
      }
    } else if (kind == "streaming") {
      // streaming <layout> <chunk rows> <streams>
      uint64_t chunk = 0;
      int streams = 0;
      ls >> l >> chunk >> streams;
      if (valid_layout(l) && chunk && chunk % 32 == 0 && streams >= 1 && streams <= 16) {
        ctx->stream_chunk_rows[l] = chunk;
        ctx->stream_count[l] = streams;
      }
    } else if (kind == "best") {
      ls >> l >> p;
      if (valid_layout(l) && p >= 0 && p <= Q1_PLACE_GLOBAL) ctx->best[l] = (q1_placement)p;
    } else if (kind == "batch") {
      // batch <layout> <placement> <blocks per compute unit> <rows per lane>
      uint32_t per_cu = 0, lane_rows = 0;
      ls >> l >> p >> per_cu >> lane_rows;
      const bool shape_ok = per_cu >= 1 && per_cu <= 64 && lane_rows && lane_rows % 16 == 0 && lane_rows <= 16384;
      if (valid_layout(l) && p >= 0 && p <= Q1_PLACE_GLOBAL && shape_ok) {
        ctx->batch_tuning[l][p].num_blocks = per_cu;
        ctx->batch_tuning[l][p].tuples_per_thread = lane_rows;
      }
    }
  }
}

// Frees a table with its context's lock held (every entry point that creates tables
// holds it when a creation fails).
static void release_table(q1_table* t) {
  if (q1_ctx* ctx = t->ctx) {
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (void* p : t->owned)
      if (p) (void)hipFree(p);
    ctx->tables.erase(t);
  }
  delete t;
}

}  // namespace q1

using namespace q1;

#define GUARD_BEGIN try {
#define GUARD_END(ctx)                                         \
  }                                                            \
  catch (const std::bad_alloc&) {                              \
    return q1::fail(ctx, Q1_ENOMEM, "host allocation failed"); \
  }                                                            \
  catch (const std::exception& ex) {                           \
    return q1::fail(ctx, Q1_EINVAL, ex.what());                \
  }

extern "C" {

const char* q1_strerror(q1_status s) {
  switch (s) {
    case Q1_OK: return "ok";
    case Q1_EINVAL: return "invalid argument";
    case Q1_ENOMEM: return "out of memory";
    case Q1_EDEVICE: return "device error";
    case Q1_EOVERFLOW: return "overflow";
    case Q1_EBUSY: return "busy";
    case Q1_ETIMEOUT: return "timed out";
    case Q1_ECANCELED: return "canceled";
  }
  return "unknown status";
}

int q1_abi_version(void) { return Q1_ABI_VERSION; }

const char* q1_last_error(q1_ctx* ctx) { return ctx ? ctx->error.c_str() : ""; }

q1_status q1_init(int device, q1_ctx** out) {
  if (!out) return Q1_EINVAL;
  *out = nullptr;
  GUARD_BEGIN
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return Q1_EDEVICE;
  if (device < 0 || device >= count) return Q1_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return Q1_EDEVICE;
  q1_ctx* ctx = new q1_ctx();
  ctx->device = device;
  for (int l = 0; l < 4; l++)
    for (int p = 0; p < 5; p++) ctx->tuning[l][p] = q1_launch{256, 1024, 0, 0};
  hipDeviceProp_t prop;
  bool ok = hipGetDeviceProperties(&prop, device) == hipSuccess;
  ctx->compute_units = ok ? prop.multiProcessorCount : 0;
  ok = ok && hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) == hipSuccess;
  ok = ok && hipEventCreate(&ctx->ev_begin) == hipSuccess && hipEventCreate(&ctx->ev_end) == hipSuccess;
  ok = ok && hipEventCreate(&ctx->ev_h2d_begin) == hipSuccess && hipEventCreate(&ctx->ev_h2d_end) == hipSuccess;
  ok = ok && hipMalloc(&ctx->d_aggr, sizeof(q1_device_aggregates)) == hipSuccess;
  ok = ok && hipHostMalloc(&ctx->h_aggr, sizeof(q1_device_aggregates), hipHostMallocDefault) == hipSuccess;
  ok = ok && hipMalloc(&ctx->d_stats, 8 * sizeof(int64_t)) == hipSuccess;
  ok = ok && hipMalloc(&ctx->spread.tables, kSpreadTables * sizeof(q1_device_aggregates)) == hipSuccess;
  ok = ok && hipMemset(ctx->spread.tables, 0, kSpreadTables * sizeof(q1_device_aggregates)) == hipSuccess;
  // the scratch of q1_enqueue_noalloc on this device (once per device and process)
  ok = ok && noalloc_setup(device, ctx->stream) == Q1_OK;
  ok = ok && hipDeviceSynchronize() == hipSuccess;
  if (!ok) {
    q1_shutdown(ctx);
    return Q1_EDEVICE;
  }
  load_tuning_file(ctx);
  *out = ctx;
  return Q1_OK;
  GUARD_END(nullptr)
}

void q1_shutdown(q1_ctx* ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  // Tables still alive lose their device memory; their handles stay valid for
  // q1_table_free only.
  for (q1_table* t : ctx->tables) {
    for (void*& p : t->owned) {
      if (p) (void)hipFree(p);
      p = nullptr;
    }
    t->ctx = nullptr;
    t->dev = q1_columns{};
    t->rows = 0;
  }
  ctx->tables.clear();
  if (ctx->batch) batch_destroy(ctx->batch);
  stream_state_release(ctx);
  if (ctx->workspace) (void)hipFree(ctx->workspace);
  if (ctx->d_stats) (void)hipFree(ctx->d_stats);
  if (ctx->spread.tables) (void)hipFree(ctx->spread.tables);
  for (auto& kv : ctx->stream_spreads) {
    // the caller's stream may still run the last launch that used them
    (void)hipEventSynchronize(kv.second.done);
    (void)hipFree(kv.second.spread.tables);
    (void)hipEventDestroy(kv.second.done);
  }
  ctx->stream_spreads.clear();
  if (ctx->h_aggr) (void)hipHostFree(ctx->h_aggr);
  if (ctx->d_aggr) (void)hipFree(ctx->d_aggr);
  for (hipEvent_t e : {ctx->ev_begin, ctx->ev_end, ctx->ev_h2d_begin, ctx->ev_h2d_end})
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

q1_status q1_device_info_get(q1_ctx* ctx, q1_device_info* out) {
  if (!ctx || !out) return Q1_EINVAL;
  hipDeviceProp_t p;
  Q1_HIP(ctx, hipGetDeviceProperties(&p, ctx->device));
  std::memset(out, 0, sizeof(*out));
  std::snprintf(out->name, sizeof(out->name), "%.63s", p.name);
  std::snprintf(out->arch, sizeof(out->arch), "%.31s", p.gcnArchName);
  out->compute_units = p.multiProcessorCount;
  out->clock_khz = p.clockRate;
  out->memory_clock_khz = p.memoryClockRate;
  out->memory_bus_width = p.memoryBusWidth;
  out->total_memory = p.totalGlobalMem;
  out->lds_per_block = p.sharedMemPerBlock;
  out->wavefront_size = p.warpSize;
  out->max_threads_per_block = p.maxThreadsPerBlock;
  return Q1_OK;
}

q1_status q1_synchronize(q1_ctx* ctx) {
  if (!ctx) return Q1_EINVAL;
  Q1_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return Q1_OK;
}

void* q1_stream(q1_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

q1_status q1_layout_column_bytes(q1_layout layout, uint64_t rows, uint64_t bytes_out[8]) {
  if (!valid_layout(layout) || !bytes_out) return Q1_EINVAL;
  column_bytes(layout, rows, bytes_out);
  return Q1_OK;
}

// ---------------------------------------------------------------------------------------
// Tables
// ---------------------------------------------------------------------------------------
static q1_status alloc_table(q1_ctx* ctx, q1_layout layout, uint64_t rows, q1_table** out) {
  uint64_t b[8];
  column_bytes(layout, rows, b);
  q1_table* t = new q1_table();
  t->ctx = ctx;
  t->layout = layout;
  t->rows = rows;
  ctx->tables.insert(t);
  for (int i = 0; i < 8; i++) {
    if (!b[i]) continue;
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, b[i]);
    if (e != hipSuccess) {
      release_table(t);
      return fail(ctx, Q1_ENOMEM, "device allocation of a table column failed");
    }
    t->owned[i] = p;
    col(t->dev, i) = p;
    t->device_bytes += b[i];
  }
  *out = t;
  return Q1_OK;
}

q1_status q1_table_upload(q1_ctx* ctx, q1_layout layout, const q1_columns* host, uint64_t rows,
                          q1_table** out) {
  if (!ctx || !host || !out || !valid_layout(layout)) return fail(ctx, Q1_EINVAL, "bad argument");
  *out = nullptr;
  GUARD_BEGIN
  std::lock_guard<std::mutex> g(ctx->lock);
  Q1_HIP(ctx, hipSetDevice(ctx->device));
  uint64_t b[8];
  column_bytes(layout, rows, b);
  for (int i = 0; i < 8; i++)
    if (b[i] && !colc(*host, i)) return fail(ctx, Q1_EINVAL, "a column required by the layout is NULL");
  q1_table* t = nullptr;
  q1_status s = alloc_table(ctx, layout, rows, &t);
  if (s != Q1_OK) return s;
  const auto w0 = std::chrono::steady_clock::now();
  bool ok = hipEventRecord(ctx->ev_h2d_begin, ctx->stream) == hipSuccess;
  for (int i = 0; i < 8 && ok; i++)
    if (b[i])
      ok = hipMemcpyAsync((void*)colc(t->dev, i), colc(*host, i), b[i], hipMemcpyHostToDevice, ctx->stream) ==
           hipSuccess;
  ok = ok && hipEventRecord(ctx->ev_h2d_end, ctx->stream) == hipSuccess;
  ok = hipStreamSynchronize(ctx->stream) == hipSuccess && ok;
  if (!ok) {
    release_table(t);
    return fail(ctx, Q1_EDEVICE, "upload of the table columns failed");
  }
  float ms = 0;
  (void)hipEventElapsedTime(&ms, ctx->ev_h2d_begin, ctx->ev_h2d_end);
  ctx->last = q1_timing{};
  ctx->last.h2d_ms = ms;
  ctx->last.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
  ctx->last.bytes_read = t->device_bytes;
  if (layout == Q1_LAYOUT_COMPRESSED_PUSHDOWN) {
    // The domain of a pushdown table is that of its compressed columns.
    q1_table tmp = *t;
    tmp.layout = Q1_LAYOUT_COMPRESSED;
    s = measure_domain(ctx, &tmp);
    t->domain = tmp.domain;
    t->fast = tmp.fast;
  } else {
    s = measure_domain(ctx, t);
  }
  if (s != Q1_OK) {
    release_table(t);
    return s;
  }
  *out = t;
  return Q1_OK;
  GUARD_END(ctx)
}

q1_status q1_table_wrap(q1_ctx* ctx, q1_layout layout, const q1_columns* device, uint64_t rows,
                        q1_table** out) {
  if (!ctx || !device || !out || !valid_layout(layout)) return fail(ctx, Q1_EINVAL, "bad argument");
  *out = nullptr;
  GUARD_BEGIN
  std::lock_guard<std::mutex> g(ctx->lock);
  uint64_t b[8];
  column_bytes(layout, rows, b);
  for (int i = 0; i < 8; i++) {
    if (!b[i]) continue;
    const void* p = colc(*device, i);
    if (!p) return fail(ctx, Q1_EINVAL, "a column required by the layout is NULL");
    if ((uintptr_t)p % 16) return fail(ctx, Q1_EINVAL, "device columns must be 16-byte aligned");
  }
  q1_table* t = new q1_table();
  t->ctx = ctx;
  t->layout = layout;
  t->rows = rows;
  t->dev = *device;
  for (int i = 0; i < 8; i++)
    if (!b[i]) col(t->dev, i) = nullptr;
  q1_table tmp = *t;
  if (layout == Q1_LAYOUT_COMPRESSED_PUSHDOWN) tmp.layout = Q1_LAYOUT_COMPRESSED;
  const q1_status s = measure_domain(ctx, &tmp);
  if (s != Q1_OK) {
    delete t;
    return s;
  }
  t->domain = tmp.domain;
  t->fast = tmp.fast;
  ctx->tables.insert(t);
  *out = t;
  return Q1_OK;
  GUARD_END(ctx)
}

q1_status q1_table_refresh_domain(q1_ctx* ctx, q1_table* t) {
  if (!ctx || !t || t->ctx != ctx) return fail(ctx, Q1_EINVAL, "bad argument");
  GUARD_BEGIN
  std::lock_guard<std::mutex> g(ctx->lock);
  q1_table tmp = *t;
  if (t->layout == Q1_LAYOUT_COMPRESSED_PUSHDOWN) tmp.layout = Q1_LAYOUT_COMPRESSED;
  const q1_status s = measure_domain(ctx, &tmp);
  if (s != Q1_OK) return s;
  t->domain = tmp.domain;
  t->fast = tmp.fast;
  return Q1_OK;
  GUARD_END(ctx)
}

q1_status q1_table_generate(q1_ctx* ctx, q1_layout layout, uint64_t seed, uint64_t rows, q1_table** out) {
  if (!ctx || !out || !valid_layout(layout) || layout == Q1_LAYOUT_COMPRESSED_PUSHDOWN)
    return fail(ctx, Q1_EINVAL, "bad argument (pushdown tables come from q1_build_filter)");
  *out = nullptr;
  GUARD_BEGIN
  std::lock_guard<std::mutex> g(ctx->lock);
  Q1_HIP(ctx, hipSetDevice(ctx->device));
  q1_table* t = nullptr;
  q1_status s = alloc_table(ctx, layout, rows, &t);
  if (s != Q1_OK) return s;
  s = launch_generate(layout, seed, rows, t->dev, ctx->stream);
  if (s == Q1_OK) s = measure_domain(ctx, t);
  if (s != Q1_OK) {
    release_table(t);
    return fail(ctx, s, "generator failed");
  }
  *out = t;
  return Q1_OK;
  GUARD_END(ctx)
}

q1_status q1_table_info(const q1_table* t, q1_layout* layout, uint64_t* rows, q1_columns* device) {
  if (!t) return Q1_EINVAL;
  if (layout) *layout = t->layout;
  if (rows) *rows = t->rows;
  if (device) *device = t->dev;
  return Q1_OK;
}

q1_status q1_table_download(q1_ctx* ctx, const q1_table* t, const q1_columns* host_out) {
  if (!ctx || !t || !host_out) return fail(ctx, Q1_EINVAL, "bad argument");
  GUARD_BEGIN
  std::lock_guard<std::mutex> g(ctx->lock);
  uint64_t b[8];
  column_bytes(t->layout, t->rows, b);
  for (int i = 0; i < 8; i++) {
    if (!b[i] || !colc(*host_out, i)) continue;
    Q1_HIP(ctx, hipMemcpyAsync((void*)colc(*host_out, i), colc(t->dev, i), b[i], hipMemcpyDeviceToHost, ctx->stream));
  }
  Q1_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return Q1_OK;
  GUARD_END(ctx)
}

uint64_t q1_table_device_bytes(const q1_table* t) { return t ? t->device_bytes : 0; }

// A table freed after its context was shut down has already lost its device memory
// (q1_shutdown releases it); only the handle is deleted then.
void q1_table_free(q1_table* t) {
  if (!t) return;
  if (q1_ctx* ctx = t->ctx) {
    std::lock_guard<std::mutex> g(ctx->lock);
    release_table(t);
  } else {
    delete t;
  }
}

q1_status q1_build_filter(q1_ctx* ctx, const q1_table* compressed, int32_t cutoff_day, q1_table** out) {
  if (!ctx || !compressed || !out || compressed->layout != Q1_LAYOUT_COMPRESSED)
    return fail(ctx, Q1_EINVAL, "q1_build_filter needs a COMPRESSED table");
  *out = nullptr;
  GUARD_BEGIN
  std::lock_guard<std::mutex> g(ctx->lock);
  q1_table* t = new q1_table();
  t->ctx = ctx;
  t->layout = Q1_LAYOUT_COMPRESSED_PUSHDOWN;
  t->rows = compressed->rows;
  t->dev = compressed->dev;
  t->dev.ship_date = nullptr;
  t->domain = compressed->domain;
  t->fast = compressed->fast;
  t->filter_cutoff = cutoff_day;
  ctx->tables.insert(t);
  const uint64_t words = (t->rows + 31) / 32;
  if (words) {
    void* f = nullptr;
    if (hipMalloc(&f, words * 4) != hipSuccess) {
      ctx->tables.erase(t);
      delete t;
      return fail(ctx, Q1_ENOMEM, "filter allocation failed");
    }
    t->owned[7] = f;
    t->dev.filter = f;
    t->device_bytes = words * 4;
    q1_status s = hipEventRecord(ctx->ev_begin, ctx->stream) == hipSuccess ? Q1_OK : Q1_EDEVICE;
    if (s == Q1_OK)
      s = launch_build_filter((const uint16_t*)compressed->dev.ship_date, t->rows, cutoff_day, (uint32_t*)f,
                              ctx->stream);
    if (s == Q1_OK && hipEventRecord(ctx->ev_end, ctx->stream) != hipSuccess) s = Q1_EDEVICE;
    if (s == Q1_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) s = Q1_EDEVICE;
    if (s != Q1_OK) {
      release_table(t);
      return fail(ctx, s, "building the filter bitmap failed");
    }
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end);
    ctx->last = q1_timing{};
    ctx->last.kernel_ms = ms;
    ctx->last.rows = t->rows;
    ctx->last.bytes_read = 2 * t->rows + 4 * words;
  }
  *out = t;
  return Q1_OK;
  GUARD_END(ctx)
}

// ---------------------------------------------------------------------------------------
// Results
// ---------------------------------------------------------------------------------------
void q1_aggregates_to_result(const q1_device_aggregates* a, q1_result* out) {
  if (!a || !out) return;
  for (int g = 0; g < Q1_NUM_GROUPS; g++) {
    q1_group& r = out->g[g];
    r.sum_qty = (int64_t)a->sum_qty[g];
    r.sum_base_price = (int64_t)a->sum_base_price[g];
    r.sum_disc_price_lo = a->sum_disc_price_lo[g];
    r.sum_disc_price_hi = (int64_t)a->sum_disc_price_hi[g];
    r.sum_charge_lo = a->sum_charge_lo[g];
    r.sum_charge_hi = (int64_t)a->sum_charge_hi[g];
    r.sum_disc = (int64_t)a->sum_disc[g];
    r.count = (int64_t)a->count[g];
  }
}

void q1_result_merge(const q1_result* a, const q1_result* b, q1_result* out) {
  if (!a || !b || !out) return;
  q1_result r;
  for (int g = 0; g < Q1_NUM_GROUPS; g++) {
    const q1_group &x = a->g[g], &y = b->g[g];
    q1_group& z = r.g[g];
    z.sum_qty = (int64_t)((uint64_t)x.sum_qty + (uint64_t)y.sum_qty);
    z.sum_base_price = (int64_t)((uint64_t)x.sum_base_price + (uint64_t)y.sum_base_price);
    z.sum_disc = (int64_t)((uint64_t)x.sum_disc + (uint64_t)y.sum_disc);
    z.count = (int64_t)((uint64_t)x.count + (uint64_t)y.count);
    z.sum_disc_price_lo = x.sum_disc_price_lo + y.sum_disc_price_lo;
    z.sum_disc_price_hi = (int64_t)((uint64_t)x.sum_disc_price_hi + (uint64_t)y.sum_disc_price_hi +
                                    (z.sum_disc_price_lo < x.sum_disc_price_lo));
    z.sum_charge_lo = x.sum_charge_lo + y.sum_charge_lo;
    z.sum_charge_hi = (int64_t)((uint64_t)x.sum_charge_hi + (uint64_t)y.sum_charge_hi +
                                (z.sum_charge_lo < x.sum_charge_lo));
  }
  *out = r;
}

// ---------------------------------------------------------------------------------------
// The query
// ---------------------------------------------------------------------------------------
static q1_status run_locked(q1_ctx* ctx, const q1_table* t, uint64_t row_begin, uint64_t row_count,
                            int32_t cutoff_day, q1_placement placement, const q1_launch* launch,
                            q1_result* out) {
  if (row_begin > t->rows || row_count > t->rows - row_begin)
    return fail(ctx, Q1_EINVAL, "row range outside the table");
  const q1_placement p = resolve_placement(ctx, t->layout, placement);
  q1_launch l;
  resolve_launch(ctx, t->layout, p, row_count, launch, &l);
  LaunchShape shape;
  const char* why = "";
  q1_status s = make_shape(p, t->layout, &l, row_begin, row_begin + row_count, &shape, &why);
  if (s != Q1_OK) return fail(ctx, s, why);
  s = ensure_workspace(ctx, scan_workspace_bytes(p, shape));
  if (s != Q1_OK) return s;
  ScanArgs a;
  a.cols = t->dev;
  a.row_begin = row_begin;
  a.row_end = row_begin + row_count;
  a.cutoff_day = t->layout == Q1_LAYOUT_COMPRESSED_PUSHDOWN ? t->filter_cutoff : cutoff_day;
  a.out = ctx->d_aggr;
  const auto w0 = std::chrono::steady_clock::now();
  s = launch_zero(ctx->d_aggr, 1, ctx->stream);
  if (s != Q1_OK) return fail(ctx, s, "zero kernel launch failed");
  Q1_HIP(ctx, hipEventRecord(ctx->ev_begin, ctx->stream));
  s = scan_on(ctx, p, t->layout, t->fast, a, shape, ctx->stream, !t->fast);
  if (s != Q1_OK) return fail(ctx, s, "scan kernel launch failed");
  Q1_HIP(ctx, hipEventRecord(ctx->ev_end, ctx->stream));
  Q1_HIP(ctx, hipMemcpyAsync(ctx->h_aggr, ctx->d_aggr, sizeof(q1_device_aggregates), hipMemcpyDeviceToHost,
                             ctx->stream));
  Q1_HIP(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end);
  q1_aggregates_to_result(ctx->h_aggr, out);
  ctx->last = q1_timing{};
  ctx->last.kernel_ms = ms;
  ctx->last.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
  ctx->last.rows = row_count;
  ctx->last.bytes_read = algorithmic_bytes(t->layout, row_count);
  ctx->last.num_blocks = shape.blocks;
  ctx->last.threads_per_block = shape.threads;
  ctx->last.tuples_per_thread = l.tuples_per_thread;
  ctx->last.placement = t->fast ? (uint32_t)p : (uint32_t)Q1_PLACE_IN_REGISTERS_PER_THREAD;
  return Q1_OK;
}

q1_status q1_run(q1_ctx* ctx, const q1_table* t, uint64_t row_begin, uint64_t row_count, int32_t cutoff_day,
                 q1_placement placement, const q1_launch* launch, q1_result* out) {
  if (!ctx || !t || !out || !valid_placement(placement)) return fail(ctx, Q1_EINVAL, "bad argument");
  GUARD_BEGIN
  std::lock_guard<std::mutex> g(ctx->lock);
  return run_locked(ctx, t, row_begin, row_count, cutoff_day, placement, launch, out);
  GUARD_END(ctx)
}

q1_status q1_last_timing(q1_ctx* ctx, q1_timing* out) {
  if (!ctx || !out) return Q1_EINVAL;
  *out = ctx->last;
  return Q1_OK;
}

q1_status q1_enqueue_noalloc(q1_ctx* ctx, q1_placement placement, q1_layout layout, const q1_launch* launch,
                             q1_device_aggregates* aggregates, const q1_columns* device, uint64_t row_begin,
                             uint64_t num_tuples, int32_t cutoff_day, void* hip_stream) {
  if (!ctx || !aggregates || !device || !valid_layout(layout) || !valid_placement(placement))
    return fail(ctx, Q1_EINVAL, "bad argument");
  GUARD_BEGIN
  std::lock_guard<std::mutex> g(ctx->lock);
  uint64_t b[8];
  column_bytes(layout, row_begin + num_tuples, b);
  for (int i = 0; i < 8; i++) {
    if (!b[i]) continue;
    const void* p = colc(*device, i);
    if (!p) return fail(ctx, Q1_EINVAL, "a column required by the layout is NULL");
    if ((uintptr_t)p % 16) return fail(ctx, Q1_EINVAL, "device columns must be 16-byte aligned");
  }
  const q1_placement p = resolve_placement(ctx, layout, placement);
  q1_launch l;
  resolve_launch(ctx, layout, p, num_tuples, launch, &l, true);
  LaunchShape shape;
  const char* why = "";
  q1_status s = make_shape(p, layout, &l, row_begin, row_begin + num_tuples, &shape, &why);
  if (s != Q1_OK) return fail(ctx, s, why);
  ScanArgs a;
  a.cols = *device;
  a.row_begin = row_begin;
  a.row_end = row_begin + num_tuples;
  a.cutoff_day = cutoff_day;
  a.out = aggregates;
  // the device's static scratch array instead of the context's spread tables and workspace
  s = noalloc_launch_scan(ctx->device, p, layout, (l.flags & Q1_LAUNCH_ASSUME_TPCH_DOMAIN) != 0, a, shape,
                          hip_stream ? hip_stream : (void*)ctx->stream, &why);
  return s == Q1_OK ? Q1_OK : fail(ctx, s, why);
  GUARD_END(ctx)
}

q1_status q1_enqueue(q1_ctx* ctx, q1_placement placement, q1_layout layout, const q1_launch* launch,
                     q1_device_aggregates* aggregates, const q1_columns* device, uint64_t row_begin,
                     uint64_t num_tuples, int32_t cutoff_day, void* hip_stream) {
  if (!ctx || !aggregates || !device || !valid_layout(layout) || !valid_placement(placement))
    return fail(ctx, Q1_EINVAL, "bad argument");
  GUARD_BEGIN
  std::lock_guard<std::mutex> g(ctx->lock);
  uint64_t b[8];
  column_bytes(layout, row_begin + num_tuples, b);
  for (int i = 0; i < 8; i++) {
    if (!b[i]) continue;
    const void* p = colc(*device, i);
    if (!p) return fail(ctx, Q1_EINVAL, "a column required by the layout is NULL");
    if ((uintptr_t)p % 16) return fail(ctx, Q1_EINVAL, "device columns must be 16-byte aligned");
  }
  const q1_placement p = resolve_placement(ctx, layout, placement);
  q1_launch l;
  resolve_launch(ctx, layout, p, num_tuples, launch, &l, true);
  LaunchShape shape;
  const char* why = "";
  q1_status s = make_shape(p, layout, &l, row_begin, row_begin + num_tuples, &shape, &why);
  if (s != Q1_OK) return fail(ctx, s, why);
  void* stream = hip_stream ? hip_stream : (void*)ctx->stream;
  // The global placement's workspace is per launch: size it and keep launches that
  // share it ordered on the context stream.
  if (p == Q1_PLACE_GLOBAL && stream != (void*)ctx->stream)
    return fail(ctx, Q1_EINVAL, "the global placement enqueues on the context stream only");
  s = ensure_workspace(ctx, scan_workspace_bytes(p, shape));
  if (s != Q1_OK) return s;
  ScanArgs a;
  a.cols = *device;
  a.row_begin = row_begin;
  a.row_end = row_begin + num_tuples;
  a.cutoff_day = cutoff_day;
  a.out = aggregates;
  const bool fast = (l.flags & Q1_LAUNCH_ASSUME_TPCH_DOMAIN) != 0;
  s = scan_on(ctx, p, layout, fast, a, shape, stream);
  return s == Q1_OK ? Q1_OK : fail(ctx, s, "scan kernel launch failed");
  GUARD_END(ctx)
}

// ---------------------------------------------------------------------------------------
// Host memory and the autotuning table
// ---------------------------------------------------------------------------------------
void* q1_host_alloc(q1_ctx* ctx, size_t bytes) {
  void* p = nullptr;
  if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
    if (ctx) ctx->error = "hipHostMalloc failed";
    return nullptr;
  }
  return p;
}

void q1_host_free(q1_ctx*, void* p) {
  if (p) (void)hipHostFree(p);
}

q1_status q1_tuning_set(q1_ctx* ctx, q1_layout layout, q1_placement placement, const q1_launch* best) {
  if (!ctx || !best || !valid_layout(layout) || placement < 0 || placement > Q1_PLACE_GLOBAL)
    return fail(ctx, Q1_EINVAL, "bad argument");
  LaunchShape shape;
  const char* why = "";
  const q1_status s = make_shape(placement, layout, best, 0, 1, &shape, &why);
  if (s != Q1_OK) return fail(ctx, s, why);
  ctx->tuning[layout][placement] = *best;
  ctx->tuning[layout][placement].num_blocks = 0;
  ctx->tuning_at[layout][placement].clear();
  return Q1_OK;
}

q1_status q1_tuning_get(q1_ctx* ctx, q1_layout layout, q1_placement placement, q1_launch* out) {
  if (!ctx || !out || !valid_layout(layout) || !valid_placement(placement)) return Q1_EINVAL;
  *out = ctx->tuning[layout][resolve_placement(ctx, layout, placement)];
  return Q1_OK;
}

q1_status q1_tuning_resolve(q1_ctx* ctx, q1_layout layout, q1_placement placement, uint64_t rows,
                            q1_launch* out) {
  if (!ctx || !out || !valid_layout(layout) || !valid_placement(placement)) return Q1_EINVAL;
  resolve_launch(ctx, layout, resolve_placement(ctx, layout, placement), rows, nullptr, out);
  return Q1_OK;
}

q1_status q1_tuning_get_streaming(q1_ctx* ctx, q1_layout layout, uint64_t* chunk_rows, int* num_streams) {
  if (!ctx || !chunk_rows || !num_streams || !valid_layout(layout)) return Q1_EINVAL;
  *chunk_rows = ctx->stream_chunk_rows[layout] ? ctx->stream_chunk_rows[layout] : kStreamChunkRows;
  *num_streams = ctx->stream_count[layout] ? ctx->stream_count[layout] : kStreamCount;
  return Q1_OK;
}

q1_status q1_tuning_set_best_placement(q1_ctx* ctx, q1_layout layout, q1_placement placement) {
  if (!ctx || !valid_layout(layout) || placement < 0 || placement > Q1_PLACE_GLOBAL) return Q1_EINVAL;
  ctx->best[layout] = placement;
  return Q1_OK;
}

int32_t q1_day_number(int32_t year, int32_t month, int32_t day) { return q1::day_number(year, month, day); }

}  // extern "C"
