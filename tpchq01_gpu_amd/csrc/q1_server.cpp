// Serving: continuous batching of Q1 requests over HBM-resident tables.
//
//   submit(): on an idle server (nothing queued, a slot free, the request admissible) the
//     submitting thread launches the request at once as a batch of one; otherwise it
//     joins the pending queue (bounded: Q1_EBUSY when full).
//   scheduler thread: whenever an in-flight slot is free, admits pending requests in
//     arrival order (optionally holding them back batch_window_us for more while other
//     batches are still running) and launches them as one batch on the slot's stream (a
//     hipGraph per slot: descriptor copy, batched kernel, and the fold of the requests'
//     spread tables, which stores the results into the slot's pinned host buffer).
//   completion: a thread per slot waits for the slot's batch and publishes every result
//     with its latency; a client waiting for a request of a launched batch waits for the
//     device itself and publishes the batch first if it gets there first.
// Large requests (at least exclusive_bytes of columns to scan) run alone, through q1_run's
// scan kernel with its tuned launch and the slot's own spread tables, and one at a time in
// arrival order: two large scans sharing the device would both finish late, which is what
// sets the latency tail.  Smaller requests batch beside them on the other slots, so a short
// request that arrives while a large one runs starts at once instead of queueing behind it.
// A batch runs the persistent batched kernel with its tuned grid, or one block per entry
// when the batch has fewer entries than that grid.
// Admission by device memory: a request pins its table's columns (counted once per table
// while any of its requests is in flight) and its descriptor and result table.
// Requests on tables outside the TPC-H value domain run alone through the 64-bit kernel.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>
#include <unordered_map>
#include <vector>

#include "q1_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Pending {
  uint64_t ticket;
  q1_request req;
  Clock::time_point submitted;
};

struct Finished {
  q1_status status = Q1_OK;
  q1_result result{};
  q1_request_stats stats{};
};

struct Where {
  size_t slot;
  uint32_t batch_id;
};

uint64_t request_state_bytes() { return sizeof(q1::BatchDesc) + sizeof(q1_device_aggregates); }

// Device bytes of a table's columns (what a request on it keeps resident).
uint64_t table_bytes(const q1_table* t) {
  uint64_t b[8] = {};
  if (q1_layout_column_bytes(t->layout, t->rows, b) != Q1_OK) return 0;
  uint64_t sum = 0;
  for (uint64_t x : b) sum += x;
  return sum;
}

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  uint32_t capacity = 0;
  q1_layout layout = Q1_LAYOUT_COMPRESSED;
  bool single = false;           // a lone request through the 64-bit scan kernel
  q1::BatchHeader* h_buf = nullptr;
  q1::BatchHeader* d_buf = nullptr;
  q1_device_aggregates* h_out = nullptr;
  q1_device_aggregates* d_out = nullptr;
  q1_device_aggregates* d_spread = nullptr;  // the batch's per-request spread tables
  uint32_t spread_requests = 0;
  q1_device_aggregates* h_out_dev = nullptr;  // h_out as the device addresses it
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraphNode_t kernel_node = nullptr;  // the batched kernel's node (its grid is set per batch)
  hipKernelNodeParams kernel_params{};
  uint32_t grid = 0;             // the batched kernel's tuned (persistent) grid
  q1::Spread spread{};           // the slot's own spread-merge tables (lone requests)
  bool spread_dirty = false;     // a failed launch may have left partial sums in them
  std::vector<Pending> batch;
  Clock::time_point started;
  uint32_t batch_id = 0;
  bool busy = false;
  bool launched = false;         // the done event has been recorded
  std::thread waiter;
};

}  // namespace

struct q1_server {
  q1_ctx* ctx = nullptr;
  q1_server_config cfg{};
  q1_placement placement = Q1_PLACE_IN_REGISTERS;
  std::mutex m;
  std::condition_variable cv_sched, cv_done, cv_slot;
  std::deque<Pending> queue;
  std::unordered_map<uint64_t, Finished> finished;
  std::unordered_map<uint64_t, Where> inflight;            // tickets on the device
  std::unordered_map<const q1_table*, uint32_t> pinned;    // in-flight requests per table
  uint64_t next_ticket = 1;
  uint64_t bytes_in_flight = 0;
  uint32_t large_in_flight = 0;                            // large requests on the device
  uint32_t batches = 0;
  bool stop = false;
  std::vector<Slot> slots;
  std::thread scheduler;
};

namespace {

void free_slot_buffers(Slot& s) {
  if (s.exec) (void)hipGraphExecDestroy(s.exec);
  if (s.graph) (void)hipGraphDestroy(s.graph);
  if (s.h_buf) (void)hipHostFree(s.h_buf);
  if (s.h_out) (void)hipHostFree(s.h_out);
  if (s.d_buf) (void)hipFree(s.d_buf);
  if (s.d_out) (void)hipFree(s.d_out);
  if (s.d_spread) (void)hipFree(s.d_spread);
  s.d_spread = nullptr;
  s.spread_requests = 0;
  s.h_out_dev = nullptr;
  s.exec = nullptr;
  s.graph = nullptr;
  s.kernel_node = nullptr;
  s.h_buf = s.d_buf = nullptr;
  s.h_out = s.d_out = nullptr;
  s.capacity = 0;
}

// (Re)builds a slot's buffers and graph for `cap` requests of `layout`.
q1_status prepare_slot(q1_server* sv, Slot& s, uint32_t cap, q1_layout layout) {
  if (s.capacity == cap && s.layout == layout && s.exec) return Q1_OK;
  free_slot_buffers(s);
  const uint64_t bb = sizeof(q1::BatchHeader) + (uint64_t)cap * sizeof(q1::BatchDesc);
  const uint64_t ob = (uint64_t)cap * sizeof(q1_device_aggregates);
  s.spread_requests = std::min(cap, q1::kBatchSpreadRequests);
  const uint64_t sb = (uint64_t)s.spread_requests * q1::kSpreadTables * sizeof(q1_device_aggregates);
  if (hipHostMalloc(&s.h_buf, bb, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc(&s.h_out, ob, hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer((void**)&s.h_out_dev, s.h_out, 0) != hipSuccess || hipMalloc(&s.d_buf, bb) != hipSuccess ||
      hipMalloc(&s.d_out, ob) != hipSuccess || hipMalloc(&s.d_spread, sb) != hipSuccess) {
    free_slot_buffers(s);
    return Q1_ENOMEM;
  }
  // zeroed once, before the capture (every fold clears the tables it read)
  if (hipMemsetAsync(s.d_spread, 0, sb, s.stream) != hipSuccess || hipStreamSynchronize(s.stream) != hipSuccess) {
    free_slot_buffers(s);
    return Q1_EDEVICE;
  }
  std::memset(s.h_buf, 0, bb);
  const uint32_t blocks = q1::batch_default_blocks(sv->ctx, layout, sv->placement);
  q1_status st = Q1_OK;
  if (hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) st = Q1_EDEVICE;
  if (st == Q1_OK && hipMemcpyAsync(s.d_buf, s.h_buf, bb, hipMemcpyHostToDevice, s.stream) != hipSuccess) st = Q1_EDEVICE;
  // With spread tables for every request the fold stores the results into h_out itself.
  const bool store = s.spread_requests == cap;
  if (st == Q1_OK && !store) st = q1::launch_zero(s.d_out, cap, s.stream);
  if (st == Q1_OK)
    st = q1::launch_batch(layout, sv->placement, s.d_buf, s.d_out, s.d_spread, s.spread_requests,
                          store ? s.h_out_dev : nullptr, blocks, q1::kBatchThreads, s.stream);
  if (st == Q1_OK && !store && hipMemcpyAsync(s.h_out, s.d_out, ob, hipMemcpyDeviceToHost, s.stream) != hipSuccess)
    st = Q1_EDEVICE;
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(s.stream, &g);
  if (st != Q1_OK || ec != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    free_slot_buffers(s);
    return Q1_EDEVICE;
  }
  s.graph = g;
  if (hipGraphInstantiate(&s.exec, g, nullptr, nullptr, 0) != hipSuccess) {
    free_slot_buffers(s);
    return Q1_EDEVICE;
  }
  // The batched kernel's node, whose grid each launch sets to the batch's entries.  If it
  // cannot be found the captured (resident) grid stays.
  size_t count = 0;
  if (hipGraphGetNodes(g, nullptr, &count) == hipSuccess && count) {
    std::vector<hipGraphNode_t> nodes(count);
    if (hipGraphGetNodes(g, nodes.data(), &count) == hipSuccess) {
      const void* fn = q1::batch_kernel_function(layout, sv->placement);
      for (hipGraphNode_t node : nodes) {
        hipGraphNodeType type;
        hipKernelNodeParams kp{};
        if (hipGraphNodeGetType(node, &type) == hipSuccess && type == hipGraphNodeTypeKernel &&
            hipGraphKernelNodeGetParams(node, &kp) == hipSuccess && kp.func == fn) {
          s.kernel_node = node;
          s.kernel_params = kp;
        }
      }
    }
  }
  s.grid = blocks;
  s.capacity = cap;
  s.layout = layout;
  return Q1_OK;
}

// Launches the slot's batch; called with the server lock released, on the server's device.
q1_status launch_slot(q1_server* sv, Slot& s) {
  const uint32_t n = (uint32_t)s.batch.size();
  const q1_table* t0 = s.batch[0].req.table;
  if (s.single) {
    // One request on a table outside the TPC-H domain: the 64-bit scan kernel.
    q1_status st = prepare_slot(sv, s, 16, t0->layout);
    if (st != Q1_OK) return st;
    const q1_request& r = s.batch[0].req;
    q1::LaunchShape shape;
    const char* why = "";
    q1_launch l{256, 1024, 0, 0};
    st = q1::make_shape(Q1_PLACE_IN_REGISTERS, t0->layout, &l, r.row_begin, r.row_begin + r.row_count, &shape, &why);
    if (st != Q1_OK) return st;
    q1::ScanArgs a;
    a.cols = t0->dev;
    a.row_begin = r.row_begin;
    a.row_end = r.row_begin + r.row_count;
    a.cutoff_day = t0->layout == Q1_LAYOUT_COMPRESSED_PUSHDOWN ? t0->filter_cutoff : r.cutoff_day;
    a.out = s.d_out;
    st = q1::launch_zero(s.d_out, 1, s.stream);
    if (st == Q1_OK)
      st = q1::launch_scan(Q1_PLACE_IN_REGISTERS, t0->layout, false, a, shape, nullptr, s.stream, nullptr, true);
    if (st == Q1_OK && hipMemcpyAsync(s.h_out, s.d_out, sizeof(q1_device_aggregates), hipMemcpyDeviceToHost,
                                      s.stream) != hipSuccess)
      st = Q1_EDEVICE;
    if (st == Q1_OK && hipEventRecord(s.done, s.stream) != hipSuccess) {
      (void)hipStreamSynchronize(s.stream);
      st = Q1_EDEVICE;
    }
    return st;
  }
  uint32_t cap = 16;
  while (cap < n) cap <<= 1;
  q1_status st = prepare_slot(sv, s, cap, t0->layout);
  if (st != Q1_OK) return st;
  if (n == 1 && s.spread.tables) {
    // A lone request runs the scan kernel of q1_run with its tuned launch (the batched
    // kernel's list of entries buys nothing for one request), merging large grids through
    // the slot's own spread tables.
    const q1_request& r = s.batch[0].req;
    q1_placement p = q1::resolve_placement(sv->ctx, t0->layout, Q1_PLACE_AUTO);
    q1_launch l;
    q1::resolve_launch(sv->ctx, t0->layout, p, r.row_count, nullptr, &l);
    q1::LaunchShape shape;
    const char* why = "";
    st = q1::make_shape(p, t0->layout, &l, r.row_begin, r.row_begin + r.row_count, &shape, &why);
    if (st == Q1_OK && q1::scan_workspace_bytes(p, shape) == 0) {
      q1::ScanArgs a;
      a.cols = t0->dev;
      a.row_begin = r.row_begin;
      a.row_end = r.row_begin + r.row_count;
      a.cutoff_day = t0->layout == Q1_LAYOUT_COMPRESSED_PUSHDOWN ? t0->filter_cutoff : r.cutoff_day;
      a.out = s.d_out;
      if (s.spread_dirty) {
        st = q1::launch_zero(s.spread.tables, q1::kSpreadTables, s.stream);
        if (st != Q1_OK) return st;
        s.spread_dirty = false;
      }
      st = q1::launch_zero(s.d_out, 1, s.stream);
      if (st == Q1_OK) st = q1::launch_scan(p, t0->layout, true, a, shape, nullptr, s.stream, &s.spread);
      if (st != Q1_OK) s.spread_dirty = true;
      if (st == Q1_OK && hipMemcpyAsync(s.h_out, s.d_out, sizeof(q1_device_aggregates), hipMemcpyDeviceToHost,
                                        s.stream) != hipSuccess)
        st = Q1_EDEVICE;
      if (st == Q1_OK && hipEventRecord(s.done, s.stream) != hipSuccess) {
        (void)hipStreamSynchronize(s.stream);
        st = Q1_EDEVICE;
      }
      return st;
    }
  }
  const uint32_t lane_rows = q1::batch_lane_rows(sv->ctx, t0->layout, sv->placement);
  q1::BatchDesc* d = reinterpret_cast<q1::BatchDesc*>(s.h_buf + 1);
  uint64_t entries = 0;
  for (uint32_t i = 0; i < n; i++) entries += q1::batch_describe(&d[i], s.batch[i].req, entries, lane_rows);
  if (entries >= (1ull << 31)) return Q1_EINVAL;
  s.h_buf->n = n;
  s.h_buf->total_blocks = entries;
  s.h_buf->tiles_per_thread = lane_rows / q1::layout_tile_rows(t0->layout);
  s.h_buf->spread_requests = s.spread_requests;
  if (s.kernel_node) {
    // The tuned persistent grid, or one block per entry for a batch with fewer entries
    // (block b then runs entry b only).
    hipKernelNodeParams kp = s.kernel_params;
    kp.gridDim = dim3((uint32_t)std::min<uint64_t>(std::max<uint64_t>(entries, 1), s.grid));
    if (hipGraphExecKernelNodeSetParams(s.exec, s.kernel_node, &kp) != hipSuccess) s.kernel_node = nullptr;
  }
  if (hipGraphLaunch(s.exec, s.stream) != hipSuccess || hipEventRecord(s.done, s.stream) != hipSuccess) {
    // The graph may still be reading the slot's pinned buffers: let it finish before the
    // caller frees the slot for the next batch, and rebuild the slot (its spread tables
    // may hold partial sums).
    (void)hipStreamSynchronize(s.stream);
    free_slot_buffers(s);
    return Q1_EDEVICE;
  }
  return Q1_OK;
}

// Device bytes admitting `r` would add to what the in-flight work pins.
uint64_t admission_cost(const q1_server* sv, const q1_request& r) {
  return request_state_bytes() + (sv->pinned.count(r.table) ? 0 : table_bytes(r.table));
}

// Whether `r` scans at least exclusive_bytes of columns (its share of the table's bytes).
bool is_large(const q1_server* sv, const q1_request& r) {
  const q1_table* t = r.table;
  if (!t->rows) return false;
  const unsigned __int128 scanned = (unsigned __int128)table_bytes(t) * r.row_count / t->rows;
  return scanned >= sv->cfg.exclusive_bytes;
}

// Whether `r` may go on the device now: it fits the memory budget and, if large, no other
// large request is in flight (server lock held).
bool admissible(const q1_server* sv, const q1_request& r) {
  if (sv->bytes_in_flight + admission_cost(sv, r) > sv->cfg.memory_budget) return false;
  return sv->large_in_flight == 0 || !is_large(sv, r);
}

// Puts a pending request into a slot's batch (server lock held).
void take(q1_server* sv, size_t slot, const Pending& p) {
  sv->bytes_in_flight += admission_cost(sv, p.req);
  if (is_large(sv, p.req)) sv->large_in_flight++;
  sv->pinned[p.req.table]++;
  sv->inflight[p.ticket] = Where{slot, sv->slots[slot].batch_id};
  sv->slots[slot].batch.push_back(p);
}

// Publishes the results of batch `id` of the slot, unless they were published already.
void publish(q1_server* sv, Slot& s, q1_status st, uint32_t id) {
  const auto now = Clock::now();
  std::lock_guard<std::mutex> g(sv->m);
  if (!s.busy || s.batch_id != id) return;
  for (size_t i = 0; i < s.batch.size(); i++) {
    Finished f;
    f.status = st;
    if (st == Q1_OK) q1_aggregates_to_result(&s.h_out[i], &f.result);
    f.stats.submit_to_start_us =
        std::chrono::duration<double, std::micro>(s.started - s.batch[i].submitted).count();
    f.stats.latency_us = std::chrono::duration<double, std::micro>(now - s.batch[i].submitted).count();
    f.stats.batch_size = (uint32_t)s.batch.size();
    f.stats.batch_id = s.batch_id;
    sv->finished[s.batch[i].ticket] = f;
    sv->inflight.erase(s.batch[i].ticket);
    const q1_table* t = s.batch[i].req.table;
    if (is_large(sv, s.batch[i].req)) sv->large_in_flight--;
    sv->bytes_in_flight -= request_state_bytes();
    if (--sv->pinned[t] == 0) {
      sv->pinned.erase(t);
      sv->bytes_in_flight -= table_bytes(t);
    }
  }
  s.batch.clear();
  s.busy = false;
  s.launched = false;
  sv->cv_done.notify_all();
  sv->cv_sched.notify_all();
}

void slot_waiter(q1_server* sv, size_t idx) {
  Slot& s = sv->slots[idx];
  (void)hipSetDevice(sv->ctx->device);
  for (;;) {
    uint32_t id;
    {
      std::unique_lock<std::mutex> lk(sv->m);
      sv->cv_slot.wait(lk, [&] { return (sv->stop && !s.busy) || (s.busy && s.launched); });
      if (!s.busy) return;
      id = s.batch_id;
    }
    const q1_status st = hipEventSynchronize(s.done) == hipSuccess ? Q1_OK : Q1_EDEVICE;
    publish(sv, s, st, id);
  }
}

// Launches the batch just put into `slot` (server lock held on entry and on return).
void launch_taken(q1_server* sv, Slot& slot, std::unique_lock<std::mutex>& lk) {
  const uint32_t id = slot.batch_id;
  lk.unlock();
  const q1_status st = launch_slot(sv, slot);
  if (st != Q1_OK) {
    publish(sv, slot, st, id);
    lk.lock();
    return;
  }
  lk.lock();
  slot.launched = true;
  sv->cv_slot.notify_all();
}

void scheduler_loop(q1_server* sv) {
  (void)hipSetDevice(sv->ctx->device);
  std::unique_lock<std::mutex> lk(sv->m);
  for (;;) {
    sv->cv_sched.wait(lk, [&] {
      if (sv->stop) return true;
      if (sv->queue.empty()) return false;
      for (auto& s : sv->slots)
        if (!s.busy) return true;
      return false;
    });
    if (sv->stop && sv->queue.empty()) return;
    if (sv->stop) {
      // drain: fail what is still queued
      while (!sv->queue.empty()) {
        Finished f;
        f.status = Q1_ECANCELED;
        sv->finished[sv->queue.front().ticket] = f;
        sv->queue.pop_front();
      }
      sv->cv_done.notify_all();
      return;
    }
    // Optionally, while other batches run, give the queue a short window to fill.
    bool others_busy = false;
    for (auto& s : sv->slots) others_busy = others_busy || s.busy;
    if (others_busy && sv->queue.size() < sv->cfg.max_batch && sv->cfg.batch_window_us) {
      sv->cv_sched.wait_for(lk, std::chrono::microseconds(sv->cfg.batch_window_us),
                            [&] { return sv->stop || sv->queue.size() >= sv->cfg.max_batch; });
      if (sv->stop) continue;
    }
    size_t idx = sv->slots.size();
    for (size_t i = 0; i < sv->slots.size(); i++)
      if (!sv->slots[i].busy) idx = i;
    if (idx == sv->slots.size() || sv->queue.empty()) continue;
    Slot* slot = &sv->slots[idx];
    slot->batch_id = ++sv->batches;
    // Admission in arrival order: the first admissible request opens the batch.  A large
    // request (or one outside the TPC-H domain) runs alone; smaller ones of the same
    // layout join it up to max_batch.  A request beyond the memory budget ends the scan, so
    // later ones cannot starve it; a large request while another is in flight, or one of
    // another layout, keeps its place in the queue.  One pass looks at no more than
    // kAdmissionScan queued requests, so a long queue of waiting large requests costs each
    // scheduling pass a bounded amount of work.
    constexpr size_t kAdmissionScan = 4096;
    std::deque<Pending> rest;
    const q1_table* t0 = nullptr;
    bool alone = false;
    for (size_t looked = 0; !sv->queue.empty() && looked < kAdmissionScan; looked++) {
      Pending p = sv->queue.front();
      sv->queue.pop_front();
      if (sv->bytes_in_flight + admission_cost(sv, p.req) > sv->cfg.memory_budget) {
        rest.push_back(p);
        break;
      }
      const bool large = is_large(sv, p.req);
      const bool joins = t0 ? !large && p.req.table->fast && p.req.table->layout == t0->layout
                            : !large || sv->large_in_flight == 0;
      if (!joins) {
        rest.push_back(p);
        continue;
      }
      if (!t0) {
        t0 = p.req.table;
        alone = large || !t0->fast;
        slot->single = !t0->fast;
      }
      take(sv, idx, p);
      if (alone || slot->batch.size() >= sv->cfg.max_batch) break;
    }
    while (!rest.empty()) {
      sv->queue.push_front(rest.back());
      rest.pop_back();
    }
    if (slot->batch.empty()) {
      // nothing admissible until in-flight memory frees or the large request in flight ends
      sv->cv_sched.wait(lk);
      continue;
    }
    slot->busy = true;
    slot->started = Clock::now();
    launch_taken(sv, *slot, lk);
  }
}

}  // namespace

extern "C" {

q1_status q1_server_create(q1_ctx* ctx, const q1_server_config* cfg, q1_server** out) {
  if (!ctx || !out) return Q1_EINVAL;
  *out = nullptr;
  q1_server_config c{256, 2, 65536, 0, 0, 512ull << 20};
  if (cfg) {
    if (cfg->max_batch) c.max_batch = cfg->max_batch;
    if (cfg->max_inflight) c.max_inflight = cfg->max_inflight;
    if (cfg->queue_capacity) c.queue_capacity = cfg->queue_capacity;
    c.batch_window_us = cfg->batch_window_us;
    c.memory_budget = cfg->memory_budget;
    if (cfg->exclusive_bytes) c.exclusive_bytes = cfg->exclusive_bytes;
  }
  if (c.max_batch > 4096 || c.max_inflight > 8) return q1::fail(ctx, Q1_EINVAL, "max_batch <= 4096, max_inflight <= 8");
  try {
    (void)hipSetDevice(ctx->device);
    if (!c.memory_budget) {
      // default: three quarters of the device's memory
      size_t free_b = 0, total_b = 0;
      if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return q1::fail(ctx, Q1_EDEVICE, "hipMemGetInfo failed");
      c.memory_budget = (uint64_t)total_b / 4 * 3;
    }
    q1_server* sv = new q1_server();
    sv->ctx = ctx;
    sv->cfg = c;
    const q1_placement best = ctx->best[Q1_LAYOUT_COMPRESSED];
    sv->placement = best == Q1_PLACE_PER_THREAD_SHARED_MEM ? best : Q1_PLACE_IN_REGISTERS;
    sv->slots.resize(c.max_inflight);
    const size_t spread_bytes = q1::kSpreadTables * sizeof(q1_device_aggregates);
    for (auto& s : sv->slots) {
      if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess ||
          hipMalloc(&s.spread.tables, spread_bytes) != hipSuccess ||
          hipMemset(s.spread.tables, 0, spread_bytes) != hipSuccess) {
        q1_server_destroy(sv);
        return q1::fail(ctx, Q1_EDEVICE, "stream or spread-table creation failed");
      }
    }
    if (hipDeviceSynchronize() != hipSuccess) {
      q1_server_destroy(sv);
      return q1::fail(ctx, Q1_EDEVICE, "device synchronization failed");
    }
    for (size_t i = 0; i < sv->slots.size(); i++) sv->slots[i].waiter = std::thread(slot_waiter, sv, i);
    sv->scheduler = std::thread(scheduler_loop, sv);
    *out = sv;
    return Q1_OK;
  } catch (const std::exception& e) {
    return q1::fail(ctx, Q1_ENOMEM, e.what());
  }
}

q1_status q1_server_submit(q1_server* sv, const q1_request* req, uint64_t* ticket) {
  if (!sv || !req || !ticket || !req->table) return Q1_EINVAL;
  const q1_table* t = req->table;
  if (req->row_begin > t->rows || req->row_count > t->rows - req->row_begin) return Q1_EINVAL;
  // could never be admitted: its table and state alone exceed the budget
  if (request_state_bytes() + table_bytes(t) > sv->cfg.memory_budget) return Q1_ENOMEM;
  std::unique_lock<std::mutex> lk(sv->m);
  if (sv->stop) return Q1_ECANCELED;
  if (sv->queue.size() >= sv->cfg.queue_capacity) return Q1_EBUSY;
  Pending p{sv->next_ticket++, *req, Clock::now()};
  *ticket = p.ticket;
  // An idle server launches the request from this thread, with no hand-off to the
  // scheduler: nothing is queued ahead of it, a slot is free and it is admissible.
  size_t idx = sv->slots.size();
  if (sv->queue.empty())
    for (size_t i = 0; i < sv->slots.size(); i++)
      if (!sv->slots[i].busy) idx = i;
  if (idx < sv->slots.size() && admissible(sv, *req)) {
    Slot& slot = sv->slots[idx];
    slot.batch_id = ++sv->batches;
    slot.single = !t->fast;
    take(sv, idx, p);
    slot.busy = true;
    slot.started = p.submitted;
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (prev != sv->ctx->device) (void)hipSetDevice(sv->ctx->device);
    launch_taken(sv, slot, lk);
    if (prev >= 0 && prev != sv->ctx->device) (void)hipSetDevice(prev);
    return Q1_OK;   // a failed launch reaches the client with the request's result
  }
  sv->queue.push_back(p);
  sv->cv_sched.notify_all();
  return Q1_OK;
}

q1_status q1_server_wait(q1_server* sv, uint64_t ticket, int64_t timeout_us, q1_result* out, q1_request_stats* stats) {
  if (!sv) return Q1_EINVAL;
  std::unique_lock<std::mutex> lk(sv->m);
  auto ready = [&] { return sv->finished.count(ticket) != 0; };
  bool known = ready() || sv->inflight.count(ticket);
  if (!known)
    for (auto& p : sv->queue) known = known || p.ticket == ticket;
  if (!known) return Q1_EINVAL;
  const auto deadline = Clock::now() + std::chrono::microseconds(timeout_us < 0 ? 0 : timeout_us);
  while (!ready()) {
    auto it = sv->inflight.find(ticket);
    if (timeout_us < 0 && it != sv->inflight.end()) {
      Slot& s = sv->slots[it->second.slot];
      const uint32_t id = it->second.batch_id;
      if (s.busy && s.launched && s.batch_id == id) {
        // the request's batch is on the device: wait for the device here rather than for
        // the completion thread's wake-up, and publish the batch if first
        lk.unlock();
        const q1_status st = hipEventSynchronize(s.done) == hipSuccess ? Q1_OK : Q1_EDEVICE;
        publish(sv, s, st, id);
        lk.lock();
        continue;
      }
    }
    if (timeout_us < 0) {
      sv->cv_done.wait(lk);
    } else if (sv->cv_done.wait_until(lk, deadline) == std::cv_status::timeout && !ready()) {
      return Q1_ETIMEOUT;
    }
  }
  Finished f = sv->finished[ticket];
  sv->finished.erase(ticket);
  if (out) *out = f.result;
  if (stats) *stats = f.stats;
  return f.status;
}

q1_status q1_server_cancel(q1_server* sv, uint64_t ticket) {
  if (!sv) return Q1_EINVAL;
  std::lock_guard<std::mutex> g(sv->m);
  for (auto it = sv->queue.begin(); it != sv->queue.end(); ++it) {
    if (it->ticket == ticket) {
      sv->queue.erase(it);
      Finished f;
      f.status = Q1_ECANCELED;
      sv->finished[ticket] = f;
      sv->cv_done.notify_all();
      return Q1_OK;
    }
  }
  // on the device or already finished: too late to cancel; otherwise unknown
  return (sv->inflight.count(ticket) || sv->finished.count(ticket)) ? Q1_EBUSY : Q1_EINVAL;
}

q1_status q1_server_drain(q1_server* sv) {
  if (!sv) return Q1_EINVAL;
  std::unique_lock<std::mutex> lk(sv->m);
  sv->cv_done.wait(lk, [&] {
    if (!sv->queue.empty()) return false;
    for (auto& s : sv->slots)
      if (s.busy) return false;
    return true;
  });
  return Q1_OK;
}

void q1_server_destroy(q1_server* sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> g(sv->m);
    sv->stop = true;
    sv->cv_sched.notify_all();
    sv->cv_slot.notify_all();
  }
  if (sv->scheduler.joinable()) sv->scheduler.join();
  // let in-flight batches finish before the waiters exit
  for (;;) {
    bool busy = false;
    {
      std::lock_guard<std::mutex> g(sv->m);
      for (auto& s : sv->slots) busy = busy || s.busy;
      sv->cv_slot.notify_all();
    }
    if (!busy) break;
    std::this_thread::sleep_for(std::chrono::microseconds(200));
  }
  {
    std::lock_guard<std::mutex> g(sv->m);
    sv->cv_slot.notify_all();
  }
  for (auto& s : sv->slots)
    if (s.waiter.joinable()) s.waiter.join();
  for (auto& s : sv->slots) {
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    free_slot_buffers(s);
    if (s.spread.tables) (void)hipFree(s.spread.tables);
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.stream) (void)hipStreamDestroy(s.stream);
  }
  delete sv;
}

}  // extern "C"
