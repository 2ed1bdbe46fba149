// gfx950 utility kernels (zero fill, spread-table fold, generator, domain statistics,
// pushdown producer), kernel dispatch and the host-side launch helpers of the TPC-H Q1
// path.  The scan kernels themselves are in q1_scan.h, instantiated per layout.
#include "q1_scan.h"

namespace q1 {
namespace {

// ------------------------------------------------------------------------------------
// Utility kernels: zero fill, generator, domain statistics, pushdown producer.
// ------------------------------------------------------------------------------------
__global__ void zero_words(uint64_t* p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Adds the kSpreadTables spread tables into `out` and clears them.  One block of four
// wavefronts; lane k reads table k, so every load of the block is issued at once and the
// sums are wavefront butterflies.  The 48 words of a table split over the waves:
//   wave 0: sum_qty, sum_base_price       (words  0..11, 64-bit)
//   wave 1: sum_disc_price lo/hi pairs    (words 12..23, 128-bit)
//   wave 2: sum_charge lo/hi pairs        (words 24..35, 128-bit)
//   wave 3: sum_disc, count               (words 36..47, 64-bit)
// Each wave clears only the words it read, after its sums consumed them.
static_assert(kSpreadTables == 64, "the fold gives each table one lane of a wavefront");
static_assert(sizeof(q1_device_aggregates) == 48 * 8, "8 fields x 6 groups");
// Block b folds tables[64 b .. 64 b + 63] into out[b] (one block for a scan's tables, one
// per request for a batch's).  STORE: out[b] receives the sums with plain stores (every
// word, zeros included) instead of atomic adds, for a destination that holds nothing yet:
// a batch's results in pinned host memory, which the stores reach over PCIe.
template <bool STORE>
__device__ __forceinline__ void fold_tables(q1_device_aggregates* tables, q1_device_aggregates* out) {
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t* mine = reinterpret_cast<uint64_t*>(tables + (uint64_t)blockIdx.x * kSpreadTables + lane) + wave * 12;
  uint64_t* dst = reinterpret_cast<uint64_t*>(out + blockIdx.x) + wave * 12;
  uint64_t v[12];
#pragma unroll
  for (int i = 0; i < 12; i++) v[i] = mine[i];
  // After the butterflies every lane holds every sum; lane i then adds word (or pair) i,
  // so the wave's destination atomics are issued together rather than one after another.
  if (wave == 0 || wave == 3) {
    uint64_t my = 0;
#pragma unroll
    for (int i = 0; i < 12; i++) {
      const uint64_t sum = wave_sum_u64(v[i]);
      my = lane == (uint32_t)i ? sum : my;
    }
    if (STORE) {
      if (lane < 12) dst[lane] = my;
    } else if (lane < 12 && my) {
      atomicAdd((unsigned long long*)&dst[lane], (unsigned long long)my);
    }
  } else {
    uint64_t my_lo = 0, my_hi = 0;
#pragma unroll
    for (int g = 0; g < kNumGroups; g++) {
      uint64_t lo = v[g], hi = v[6 + g];
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) {
        const uint64_t olo = __shfl_xor(lo, m, 64), ohi = __shfl_xor(hi, m, 64);
        const uint64_t nlo = lo + olo;
        hi += ohi + (nlo < lo);
        lo = nlo;
      }
      my_lo = lane == (uint32_t)g ? lo : my_lo;
      my_hi = lane == (uint32_t)g ? hi : my_hi;
    }
    if (lane < (uint32_t)kNumGroups) {
      if (STORE) {
        dst[lane] = my_lo;
        dst[6 + lane] = my_hi;
      } else {
        atomic_add128(&dst[lane], &dst[6 + lane], my_lo, my_hi);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 12; i++) mine[i] = 0;
}

__global__ void __launch_bounds__(256) fold_spread(q1_device_aggregates* tables, q1_device_aggregates* out) {
  fold_tables<false>(tables, out);
}
__global__ void __launch_bounds__(256) fold_spread_store(q1_device_aggregates* tables, q1_device_aggregates* out) {
  fold_tables<true>(tables, out);
}

// One thread per 32-row group writes the rows' values and the packed words of the group.
template <int L>
__global__ void generate_rows(uint64_t seed, uint64_t rows, q1_columns c) {
  const uint64_t grp = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t r0 = grp * 32;
  if (r0 >= rows) return;
  const uint64_t r1 = r0 + 32 < rows ? r0 + 32 : rows;
  uint32_t rfw[2] = {0, 0}, lsw = 0;
  for (uint64_t r = r0; r < r1; r++) {
    const GenRow g = gen_row(seed, r);
    const uint32_t j = (uint32_t)(r - r0);
    if (L == Q1_LAYOUT_REFERENCE) {
      ((int32_t*)c.ship_date)[r] = g.ship_date;
      ((int64_t*)c.quantity)[r] = (int64_t)g.quantity * 100;
      ((int64_t*)c.extended_price)[r] = g.extended_price;
      ((int64_t*)c.discount)[r] = g.discount;
      ((int64_t*)c.tax)[r] = g.tax;
      ((char*)c.return_flag)[r] = g.return_flag;
      ((char*)c.line_status)[r] = g.line_status;
    } else if (L == Q1_LAYOUT_UNCOMPRESSED) {
      ((int32_t*)c.ship_date)[r] = g.ship_date;
      ((int32_t*)c.quantity)[r] = g.quantity * 100;
      ((int64_t*)c.extended_price)[r] = g.extended_price;
      ((int8_t*)c.discount)[r] = (int8_t)g.discount;
      ((int8_t*)c.tax)[r] = (int8_t)g.tax;
      ((char*)c.return_flag)[r] = g.return_flag;
      ((char*)c.line_status)[r] = g.line_status;
    } else {
      ((uint16_t*)c.ship_date)[r] = (uint16_t)(g.ship_date - kShipDateFrameOfReference);
      ((uint8_t*)c.quantity)[r] = (uint8_t)g.quantity;
      ((uint32_t*)c.extended_price)[r] = (uint32_t)g.extended_price;
      ((uint8_t*)c.discount)[r] = (uint8_t)g.discount;
      ((uint8_t*)c.tax)[r] = (uint8_t)g.tax;
      rfw[j >> 4] |= encode_return_flag((uint32_t)g.return_flag) << (2 * (j & 15));
      lsw |= encode_line_status((uint32_t)g.line_status) << j;
    }
  }
  if (L == Q1_LAYOUT_COMPRESSED || L == Q1_LAYOUT_COMPRESSED_PUSHDOWN) {
    uint32_t* rf = (uint32_t*)c.return_flag;
    rf[2 * grp] = rfw[0];
    if (r0 + 16 < rows) rf[2 * grp + 1] = rfw[1];
    ((uint32_t*)c.line_status)[grp] = lsw;
  }
}

// min/max of discount, tax, quantity, price: out[0..7] = {min,max} x {disc,tax,qty,price}.
template <int L>
__global__ void domain_stats(q1_columns c, uint64_t rows, int64_t* out) {
  int64_t mn[4] = {INT64_MAX, INT64_MAX, INT64_MAX, INT64_MAX};
  int64_t mx[4] = {INT64_MIN, INT64_MIN, INT64_MIN, INT64_MIN};
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < rows;
       r += (uint64_t)gridDim.x * blockDim.x) {
    int64_t v[4];
    if (L == Q1_LAYOUT_REFERENCE) {
      v[0] = ((const int64_t*)c.discount)[r];
      v[1] = ((const int64_t*)c.tax)[r];
      v[2] = ((const int64_t*)c.quantity)[r];
      v[3] = ((const int64_t*)c.extended_price)[r];
    } else if (L == Q1_LAYOUT_UNCOMPRESSED) {
      v[0] = ((const int8_t*)c.discount)[r];
      v[1] = ((const int8_t*)c.tax)[r];
      v[2] = ((const int32_t*)c.quantity)[r];
      v[3] = ((const int64_t*)c.extended_price)[r];
    } else {
      v[0] = ((const uint8_t*)c.discount)[r];
      v[1] = ((const uint8_t*)c.tax)[r];
      v[2] = ((const uint8_t*)c.quantity)[r];
      v[3] = ((const uint32_t*)c.extended_price)[r];
    }
    for (int i = 0; i < 4; i++) {
      mn[i] = v[i] < mn[i] ? v[i] : mn[i];
      mx[i] = v[i] > mx[i] ? v[i] : mx[i];
    }
  }
  for (int i = 0; i < 4; i++) {
    for (int m = 32; m >= 1; m >>= 1) {
      const int64_t omn = __shfl_xor(mn[i], m, 64), omx = __shfl_xor(mx[i], m, 64);
      mn[i] = omn < mn[i] ? omn : mn[i];
      mx[i] = omx > mx[i] ? omx : mx[i];
    }
  }
  if ((threadIdx.x & 63) == 0) {
    for (int i = 0; i < 4; i++) {
      atomicMin((long long*)&out[2 * i], (long long)mn[i]);
      atomicMax((long long*)&out[2 * i + 1], (long long)mx[i]);
    }
  }
}

// The pushdown producer (src/main.cu:362-389 semantics, runtime cutoff).  A lane turns the
// 8 ship dates of rows [8g, 8g + 8) -- one 16-byte load -- into 8 filter bits, and four
// neighbouring lanes' bytes make one 32-bit word of the bitmap, which the first of them
// writes.  Each lane takes kFilterGroups such groups a block-width apart (a wavefront
// reads 1 KiB of the column contiguously per group) and issues all their loads before it
// uses any: a block covers 8,192 rows, a quarter of the blocks of one lane per group, and
// each lane keeps 64 bytes in flight.  At 600M rows 4 groups per lane took 0.200-0.209 ms,
// 1, 8, 16 and 32 groups 0.232, 0.211, 0.218 and 0.228 ms.  The column must be 16-byte
// aligned (as every table column is).
constexpr int kFilterGroups = 4;
__global__ void __launch_bounds__(256) build_filter(const uint16_t* ship, uint64_t rows, int32_t cut,
                                                    uint32_t* filter) {
  const uint64_t full = rows / 8;  // groups of 8 rows inside the column
  const uint64_t g0 = (uint64_t)blockIdx.x * 256 * kFilterGroups + threadIdx.x;
  uint4 v[kFilterGroups];
#pragma unroll
  for (int k = 0; k < kFilterGroups; k++) {
    const uint64_t g = g0 + (uint64_t)k * 256;
    v[k] = g < full ? nt_load(reinterpret_cast<const uint4*>(ship) + g) : make_uint4(0, 0, 0, 0);
  }
#pragma unroll
  for (int k = 0; k < kFilterGroups; k++) {
    const uint64_t g = g0 + (uint64_t)k * 256, r0 = g * 8;
    uint32_t bits = 0;
    if (g < full) {
      const uint32_t w[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
#pragma unroll
      for (int q = 0; q < 4; q++)
        bits |= (uint32_t)((int32_t)(w[q] & 0xFFFFu) <= cut) << (2 * q) |
                (uint32_t)((int32_t)(w[q] >> 16) <= cut) << (2 * q + 1);
    } else if (r0 < rows) {
      for (uint32_t j = 0; r0 + j < rows; j++) bits |= (uint32_t)((int32_t)ship[r0 + j] <= cut) << j;
    }
    // every lane takes part in the exchange, inside the range or not
    uint32_t word = bits << (8 * (threadIdx.x & 3));
    word |= __shfl_xor(word, 1, 64);
    word |= __shfl_xor(word, 2, 64);
    if ((threadIdx.x & 3) == 0 && r0 < rows) filter[g / 4] = word;
  }
}

// ------------------------------------------------------------------------------------
// Dispatch
// ------------------------------------------------------------------------------------
FastFn fast_kernel(q1_layout l, int p, uint32_t threads) {
  switch (l) {
    case Q1_LAYOUT_REFERENCE: return fast_kernel_reference(p, threads);
    case Q1_LAYOUT_UNCOMPRESSED: return fast_kernel_uncompressed(p, threads);
    case Q1_LAYOUT_COMPRESSED: return fast_kernel_compressed(p, threads);
    default: return fast_kernel_pushdown(p, threads);
  }
}

// outside: the table is known to hold values outside the domain.
GenFn general_kernel(q1_layout l, uint32_t threads, bool outside) {
  switch (l) {
    case Q1_LAYOUT_REFERENCE: return outside ? wide_kernel_reference(threads) : any_kernel_reference(threads);
    case Q1_LAYOUT_UNCOMPRESSED: return outside ? wide_kernel_uncompressed(threads) : any_kernel_uncompressed(threads);
    case Q1_LAYOUT_COMPRESSED: return outside ? wide_kernel_compressed(threads) : any_kernel_compressed(threads);
    default: return outside ? wide_kernel_pushdown(threads) : any_kernel_pushdown(threads);
  }
}

uint32_t tile_rows(q1_layout l) {
  return l == Q1_LAYOUT_REFERENCE ? 4u : (l == Q1_LAYOUT_UNCOMPRESSED ? 8u : 16u);
}

q1_status hip_status(hipError_t e) { return e == hipSuccess ? Q1_OK : Q1_EDEVICE; }

}  // namespace

// ------------------------------------------------------------------------------------
// Host-side launch helpers
// ------------------------------------------------------------------------------------
bool domain_is_tpch(q1_layout layout, const Domain& d) {
  const int64_t max_qty =
      (layout == Q1_LAYOUT_COMPRESSED || layout == Q1_LAYOUT_COMPRESSED_PUSHDOWN) ? 63 : 6300;
  return d.min_disc >= 0 && d.max_disc <= 15 && d.min_tax >= 0 && d.max_tax <= 15 && d.min_qty >= 0 &&
         d.max_qty <= max_qty && d.min_price >= 0 && d.max_price < (1ll << 24);
}

q1_status make_shape(q1_placement placement, q1_layout layout, const q1_launch* launch,
                     uint64_t row_begin, uint64_t row_end, LaunchShape* shape, const char** why) {
  const uint32_t trows = tile_rows(layout);
  uint32_t threads = launch && launch->threads_per_block ? launch->threads_per_block : 256;
  uint32_t tpt = launch && launch->tuples_per_thread ? launch->tuples_per_thread : 1024;
  if (threads != 64 && threads != 128 && threads != 256) {
    *why = "threads_per_block must be 64, 128 or 256";
    return Q1_EINVAL;
  }
  if (tpt % 16 != 0 || tpt == 0 || tpt > 16384) {
    *why = "tuples_per_thread must be a multiple of 16 in [16, 16384]";
    return Q1_EINVAL;
  }
  if ((int)placement < 0 || placement > Q1_PLACE_GLOBAL) {
    *why = "unknown placement";
    return Q1_EINVAL;
  }
  const uint32_t tiles_per_thread = tpt / trows;
  const uint64_t first = align_rows_down(row_begin) / trows, end = (row_end + trows - 1) / trows;
  const uint64_t tiles = end > first ? end - first : 0;
  const uint64_t per_block = (uint64_t)threads * tiles_per_thread;
  uint64_t blocks = (tiles + per_block - 1) / per_block;
  if (launch && launch->num_blocks) {
    // A caller-chosen grid must still cover the range with the chosen tuples per thread.
    if ((uint64_t)launch->num_blocks < blocks) {
      *why = "num_blocks too small to cover the rows with tuples_per_thread";
      return Q1_EINVAL;
    }
    blocks = launch->num_blocks;
  }
  // HIP's grid size counts work-items in 32 bits.
  if (blocks * threads > 0xFFFFFFFFull) {
    *why = "grid too large (blocks x threads_per_block must stay below 2^32)";
    return Q1_EINVAL;
  }
  shape->blocks = (uint32_t)blocks;
  shape->threads = threads;
  shape->tiles_per_thread = tiles_per_thread;
  return Q1_OK;
}

uint64_t scan_workspace_bytes(q1_placement placement, const LaunchShape& shape) {
  return placement == Q1_PLACE_GLOBAL ? (uint64_t)shape.blocks * AccGlobal::W * kNumGroups * 8 : 0;
}

q1_status launch_scan(q1_placement placement, q1_layout layout, bool fast, const ScanArgs& args,
                      const LaunchShape& shape, void* workspace, void* stream, const Spread* spread,
                      bool table_outside) {
  if (shape.blocks == 0) return Q1_OK;
  const hipStream_t s = (hipStream_t)stream;
  const uint32_t nwaves = shape.threads / 64;
  ScanArgs a = args;
  a.out_mask = 0;
  if (fast) {
    const bool spread_merge = spread && spread->tables && shape.blocks >= kSpreadMinBlocks;
    if (spread_merge) {
      a.out = spread->tables;
      a.out_mask = kSpreadTables - 1;
    }
    const uint64_t lane_rows = (uint64_t)shape.tiles_per_thread * tile_rows(layout);
    const bool shared = placement == Q1_PLACE_PER_THREAD_SHARED_MEM;
    // The packed slot holds quantities in whole units: the compressed layouts only.
    const bool units = layout == Q1_LAYOUT_COMPRESSED || layout == Q1_LAYOUT_COMPRESSED_PUSHDOWN;
    const int variant = shared && units && lane_rows <= 256 ? kPlaceSharedPacked
                        : shared && lane_rows <= 256        ? kPlaceSharedNarrow
                                                            : (int)placement;
    uint32_t lds = variant == kPlaceSharedPacked ? packed_table_bytes(shape.threads)
                   : shared                     ? shared_table_bytes(shape.threads)
                                                : 0;
    lds += nwaves * kNumGroups * NF * 8;
    hipLaunchKernelGGL(fast_kernel(layout, variant, shape.threads), dim3(shape.blocks), dim3(shape.threads), lds,
                       s, a, shape.tiles_per_thread, (uint64_t*)workspace);
    if (spread_merge) {
      const q1_status st = hip_status(hipGetLastError());
      if (st != Q1_OK) return st;
      hipLaunchKernelGGL(fold_spread, dim3(1), dim3(256), 0, s, spread->tables, args.out);
    }
  } else {
    const uint32_t lds = shared_table_bytes(shape.threads) + nwaves * kNumGroups * 8 * 8;
    hipLaunchKernelGGL(general_kernel(layout, shape.threads, table_outside), dim3(shape.blocks), dim3(shape.threads),
                       lds, s, a, shape.tiles_per_thread);
  }
  return hip_status(hipGetLastError());
}

q1_status launch_zero(q1_device_aggregates* table, uint32_t count, void* stream) {
  const uint64_t n = (uint64_t)count * (sizeof(q1_device_aggregates) / 8);
  if (n == 0) return Q1_OK;
  hipLaunchKernelGGL(zero_words, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     (uint64_t*)table, n);
  return hip_status(hipGetLastError());
}

q1_status launch_generate(q1_layout layout, uint64_t seed, uint64_t rows, const q1_columns& cols,
                          void* stream) {
  if (rows == 0) return Q1_OK;
  const uint64_t groups = (rows + 31) / 32;
  const dim3 grid((uint32_t)((groups + 255) / 256)), block(256);
  const hipStream_t s = (hipStream_t)stream;
  switch (layout) {
    case Q1_LAYOUT_REFERENCE: hipLaunchKernelGGL(generate_rows<Q1_LAYOUT_REFERENCE>, grid, block, 0, s, seed, rows, cols); break;
    case Q1_LAYOUT_UNCOMPRESSED: hipLaunchKernelGGL(generate_rows<Q1_LAYOUT_UNCOMPRESSED>, grid, block, 0, s, seed, rows, cols); break;
    default: hipLaunchKernelGGL(generate_rows<Q1_LAYOUT_COMPRESSED>, grid, block, 0, s, seed, rows, cols); break;
  }
  return hip_status(hipGetLastError());
}

q1_status launch_domain(q1_layout layout, const q1_columns& cols, uint64_t rows, int64_t* out_dev,
                        void* stream) {
  if (rows == 0) return Q1_OK;
  uint64_t blocks = (rows + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  const dim3 grid((uint32_t)blocks), block(256);
  const hipStream_t s = (hipStream_t)stream;
  switch (layout) {
    case Q1_LAYOUT_REFERENCE: hipLaunchKernelGGL(domain_stats<Q1_LAYOUT_REFERENCE>, grid, block, 0, s, cols, rows, out_dev); break;
    case Q1_LAYOUT_UNCOMPRESSED: hipLaunchKernelGGL(domain_stats<Q1_LAYOUT_UNCOMPRESSED>, grid, block, 0, s, cols, rows, out_dev); break;
    default: hipLaunchKernelGGL(domain_stats<Q1_LAYOUT_COMPRESSED>, grid, block, 0, s, cols, rows, out_dev); break;
  }
  return hip_status(hipGetLastError());
}

q1_status launch_build_filter(const uint16_t* ship, uint64_t rows, int32_t cutoff_day, uint32_t* filter,
                              void* stream) {
  if (rows == 0) return Q1_OK;
  const int64_t c = (int64_t)cutoff_day - kShipDateFrameOfReference;
  const int32_t cut = c < -1 ? -1 : (c > 65535 ? 65535 : (int32_t)c);
  const uint64_t groups = (rows + 7) / 8;   // 8 rows per lane and group
  const uint64_t per_block = 256ull * kFilterGroups;
  hipLaunchKernelGGL(build_filter, dim3((uint32_t)((groups + per_block - 1) / per_block)), dim3(256), 0,
                     (hipStream_t)stream, ship, rows, cut, filter);
  return hip_status(hipGetLastError());
}


uint32_t layout_tile_rows(q1_layout layout) { return tile_rows(layout); }

static void* batch_kernel(q1_layout layout, q1_placement p) {
  switch (layout) {
    case Q1_LAYOUT_REFERENCE: return batch_kernel_reference(p);
    case Q1_LAYOUT_UNCOMPRESSED: return batch_kernel_uncompressed(p);
    case Q1_LAYOUT_COMPRESSED: return batch_kernel_compressed(p);
    default: return batch_kernel_pushdown(p);
  }
}

static uint32_t batch_lds(q1_layout layout, q1_placement p) {
  const bool units = layout == Q1_LAYOUT_COMPRESSED || layout == Q1_LAYOUT_COMPRESSED_PUSHDOWN;
  uint32_t lds = p != Q1_PLACE_PER_THREAD_SHARED_MEM ? 0
                 : units                            ? packed_table_bytes(kBatchBlock)
                                                    : shared_table_bytes(kBatchBlock);
  return lds + kBatchBlock / 64 * kNumGroups * NF * 8;
}

q1_placement batch_placement(q1_placement p) {
  return p == Q1_PLACE_PER_THREAD_SHARED_MEM ? p : Q1_PLACE_IN_REGISTERS;
}

uint32_t batch_max_lane_rows(q1_placement p) { return batch_placement(p) == Q1_PLACE_PER_THREAD_SHARED_MEM ? 256 : 16384; }

uint32_t batch_blocks_per_cu(q1_layout layout, q1_placement p) {
  p = batch_placement(p);
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)batch_kernel(layout, p), kBatchBlock,
                                                   batch_lds(layout, p)) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return (uint32_t)per_cu;
}

q1_status launch_batch(q1_layout layout, q1_placement placement, BatchHeader* buf, q1_device_aggregates* outs,
                       q1_device_aggregates* spread, uint32_t spread_requests, q1_device_aggregates* store_results,
                       uint32_t blocks, uint32_t threads, void* stream) {
  placement = batch_placement(placement);
  if (threads != (uint32_t)kBatchBlock || blocks == 0 || (spread_requests && !spread) ||
      (store_results && !spread_requests))
    return Q1_EINVAL;
  const BatchDesc* descs = reinterpret_cast<const BatchDesc*>(buf + 1);
  void* args[] = {(void*)&buf, (void*)&descs, (void*)&outs, (void*)&spread};
  const hipStream_t s = (hipStream_t)stream;
  q1_status st = hip_status(
      hipLaunchKernel(batch_kernel(layout, placement), dim3(blocks), dim3(threads), args, batch_lds(layout, placement), s));
  if (st == Q1_OK && spread_requests) {
    if (store_results)
      hipLaunchKernelGGL(fold_spread_store, dim3(spread_requests), dim3(256), 0, s, spread, store_results);
    else
      hipLaunchKernelGGL(fold_spread, dim3(spread_requests), dim3(256), 0, s, spread, outs);
    st = hip_status(hipGetLastError());
  }
  return st;
}

const void* batch_kernel_function(q1_layout layout, q1_placement placement) {
  return batch_kernel(layout, batch_placement(placement));
}

}  // namespace q1
