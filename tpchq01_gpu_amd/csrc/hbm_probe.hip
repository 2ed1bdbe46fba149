// hbm_probe: the practical HBM read ceiling for the Q1 scan on this GPU.
//
// Reads the nine compressed Q1 columns of N rows (u16 ship date, u8 quantity, u32 price,
// u8 discount, u8 tax, 2-bit return flags, 1-bit line statuses: 9.375 bytes per row) with
// the scan kernel's own access pattern and launch shapes, but with no filter and no
// aggregation (each lane only XORs what it loads), and then the same number of bytes as
// one flat buffer, and with the streaming pattern of the shared placement (4-row tiles,
// non-temporal loads); then the columns of the uncompressed (20 B/row) and reference
// (38 B/row) layouts with the patterns of their kernels, and a plain 2 GiB device copy
// (read + write rate), and the batched kernel's resident-grid structure over the compressed
// columns.  Each shape is timed as 20 back-to-back launches, as bench.py
// times the scan.  The fastest of these is the rate the scan could reach if its arithmetic
// were free; the scan's own rate is compared with it and with the published 8 TB/s.
//
//   hbm_probe [rows]     (default 600000000; prints one JSON line)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                  \
  do {                                                                            \
    const hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess) {                                                       \
      std::fprintf(stderr, "hbm_probe: %s failed: %s\n", #x, hipGetErrorString(e_)); \
      std::exit(1);                                                               \
    }                                                                             \
  } while (0)

namespace {

struct Columns {
  const uint4* ship;      // 2 x uint4 per 16-row tile
  const uint4* quantity;  // 1 x uint4 per tile
  const uint4* price;     // 4 x uint4 per tile
  const uint4* discount;
  const uint4* tax;
  const uint32_t* return_flag;  // 1 word per tile
  const uint16_t* line_status;  // 1 half-word per tile
};

constexpr uint32_t kSentinel = 0x9E3779B9u;  // a lane stores only if its XOR hits this

__device__ __forceinline__ uint32_t fold(uint4 v) { return v.x ^ v.y ^ v.z ^ v.w; }

// The scan kernel's pattern: in iteration k thread i reads lane tile base + k*B + i.
template <int B>
__global__ __launch_bounds__(B) void read_columns(Columns c, uint64_t tiles, uint32_t tiles_per_thread,
                                                  uint32_t* sink) {
  const uint64_t begin = (uint64_t)blockIdx.x * B * tiles_per_thread;
  const uint64_t end = begin + (uint64_t)B * tiles_per_thread < tiles ? begin + (uint64_t)B * tiles_per_thread : tiles;
  uint32_t acc = 0;
  for (uint64_t t = begin + threadIdx.x; t < end; t += B) {
    const uint4 s0 = c.ship[2 * t], s1 = c.ship[2 * t + 1];
    const uint4 q = c.quantity[t], d = c.discount[t], x = c.tax[t];
    const uint4 p0 = c.price[4 * t], p1 = c.price[4 * t + 1], p2 = c.price[4 * t + 2], p3 = c.price[4 * t + 3];
    acc += fold(s0) ^ fold(s1) ^ fold(q) ^ fold(d) ^ fold(x) ^ fold(p0) ^ fold(p1) ^ fold(p2) ^ fold(p3) ^
           c.return_flag[t] ^ c.line_status[t];
  }
  if (acc == kSentinel) sink[0] = acc;
}

// The streaming pattern: 4-row tiles (lane i of a wavefront reads tile base + i, so every
// load instruction covers whole cache lines), non-temporal loads for the five value
// columns and plain byte loads for the packed flag columns.
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));

template <int B>
__global__ __launch_bounds__(B) void read_columns_nt(Columns c, uint64_t tiles4, uint32_t tiles_per_thread,
                                                     uint32_t* sink) {
  const uint64_t begin = (uint64_t)blockIdx.x * B * tiles_per_thread;
  const uint64_t end = begin + (uint64_t)B * tiles_per_thread < tiles4 ? begin + (uint64_t)B * tiles_per_thread : tiles4;
  const uint8_t* rf = reinterpret_cast<const uint8_t*>(c.return_flag);
  const uint8_t* ls = reinterpret_cast<const uint8_t*>(c.line_status);
  uint32_t acc = 0;
  for (uint64_t t = begin + threadIdx.x; t < end; t += B) {
    const u32x2_t s = __builtin_nontemporal_load(reinterpret_cast<const u32x2_t*>(c.ship) + t);
    const uint32_t q = __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(c.quantity) + t);
    const uint32_t d = __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(c.discount) + t);
    const uint32_t x = __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(c.tax) + t);
    const u32x4_t p = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(c.price) + t);
    acc += s.x ^ s.y ^ q ^ d ^ x ^ p.x ^ p.y ^ p.z ^ p.w ^ rf[t] ^ (ls[t >> 1] >> (4 * (t & 1)));
  }
  if (acc == kSentinel) sink[0] = acc;
}

// The batched kernel's structure with the same loads and no arithmetic: a resident grid
// works through a list of entries of B lanes x tiles_per_entry 4-row tiles with a fixed
// stride, block b running entries b, b + gridDim.x, b + 2 gridDim.x, ... and waiting at a
// barrier after each entry (where the batched kernel merges).
template <int B>
__global__ __launch_bounds__(B) void read_columns_resident(Columns c, uint64_t tiles4, uint32_t tiles_per_entry,
                                                           uint32_t* sink) {
  const uint64_t per_entry = (uint64_t)B * tiles_per_entry;
  const uint64_t entries = (tiles4 + per_entry - 1) / per_entry;
  const uint8_t* rf = reinterpret_cast<const uint8_t*>(c.return_flag);
  const uint8_t* ls = reinterpret_cast<const uint8_t*>(c.line_status);
  uint32_t acc = 0;
  for (uint64_t v = blockIdx.x; v < entries; v += gridDim.x) {
    const uint64_t begin = v * per_entry;
    const uint64_t end = begin + per_entry < tiles4 ? begin + per_entry : tiles4;
    for (uint64_t t = begin + threadIdx.x; t < end; t += B) {
      const u32x2_t s = __builtin_nontemporal_load(reinterpret_cast<const u32x2_t*>(c.ship) + t);
      const uint32_t q = __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(c.quantity) + t);
      const uint32_t d = __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(c.discount) + t);
      const uint32_t x = __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(c.tax) + t);
      const u32x4_t p = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(c.price) + t);
      acc += s.x ^ s.y ^ q ^ d ^ x ^ p.x ^ p.y ^ p.z ^ p.w ^ rf[t] ^ (ls[t >> 1] >> (4 * (t & 1)));
    }
    __syncthreads();
  }
  if (acc == kSentinel) sink[0] = acc;
}

// A plain device copy, 16-byte loads and stores, each block a contiguous range: the
// read-plus-write rate a stream copy reaches on this GPU.
template <int B>
__global__ __launch_bounds__(B) void copy_flat(const uint4* src, uint4* dst, uint64_t n, uint32_t per_thread) {
  const uint64_t begin = (uint64_t)blockIdx.x * B * per_thread;
  const uint64_t end = begin + (uint64_t)B * per_thread < n ? begin + (uint64_t)B * per_thread : n;
  for (uint64_t i = begin + threadIdx.x; i < end; i += B) dst[i] = src[i];
}

// The wide layouts' columns: lane tiles of TILE rows, column k holding W[k] bytes per
// tile.  With TILE = 2 every column load of a wavefront is one contiguous stretch with at
// most 16 bytes per lane, loaded non-temporally (the pattern of the LDS placement's
// kernels); with TILE = 8 (uncompressed) or 4 (reference) and ordinary loads it is the
// pattern of the register placements.
template <int W, bool NT>
__device__ __forceinline__ uint32_t load_bytes(const char* col, uint64_t t) {
  const char* p = col + t * W;
  if constexpr (W == 2) {
    return NT ? __builtin_nontemporal_load(reinterpret_cast<const uint16_t*>(p)) : *reinterpret_cast<const uint16_t*>(p);
  } else if constexpr (W == 4) {
    return NT ? __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(p)) : *reinterpret_cast<const uint32_t*>(p);
  } else if constexpr (W == 8) {
    const u32x2_t v = NT ? __builtin_nontemporal_load(reinterpret_cast<const u32x2_t*>(p))
                         : *reinterpret_cast<const u32x2_t*>(p);
    return v.x ^ v.y;
  } else {
    uint32_t a = 0;
#pragma unroll
    for (int k = 0; k < W / 16; k++) {
      const u32x4_t v = NT ? __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(p) + k)
                           : reinterpret_cast<const u32x4_t*>(p)[k];
      a ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    return a;
  }
}

struct WideColumns {
  const char* col[7];
};

template <bool NT, int... W>
struct WideTile {
  static constexpr int kColumns = sizeof...(W);
  static constexpr int kBytes = (W + ...);
  static __device__ __forceinline__ uint32_t load(const WideColumns& c, uint64_t t) {
    int i = 0;
    uint32_t a = 0;
    ((a ^= load_bytes<W, NT>(c.col[i++], t)), ...);
    return a;
  }
};

// G tiles per lane in flight: lane i of the block reads tiles t0 + i + m*B, m < G.
template <int B, int G, typename T>
__global__ __launch_bounds__(B) void read_wide(WideColumns c, uint64_t tiles, uint32_t tiles_per_thread,
                                               uint32_t* sink) {
  const uint64_t begin = (uint64_t)blockIdx.x * B * tiles_per_thread;
  const uint64_t end = begin + (uint64_t)B * tiles_per_thread < tiles ? begin + (uint64_t)B * tiles_per_thread : tiles;
  uint32_t acc = 0;
  for (uint32_t k = 0; k < tiles_per_thread; k += G) {
    const uint64_t t0 = begin + (uint64_t)k * B + threadIdx.x;
    if (t0 + (uint64_t)(G - 1) * B < end) {
      uint32_t v[G];
#pragma unroll
      for (int m = 0; m < G; m++) v[m] = T::load(c, t0 + (uint64_t)m * B);
#pragma unroll
      for (int m = 0; m < G; m++) acc += v[m];
    } else {
      for (int m = 0; m < G; m++)
        if (t0 + (uint64_t)m * B < end) acc += T::load(c, t0 + (uint64_t)m * B);
    }
  }
  if (acc == kSentinel) sink[0] = acc;
}

// One flat buffer, 16-byte loads, each block a contiguous range.
template <int B>
__global__ __launch_bounds__(B) void read_flat(const uint4* a, uint64_t n, uint32_t per_thread, uint32_t* sink) {
  const uint64_t begin = (uint64_t)blockIdx.x * B * per_thread;
  const uint64_t end = begin + (uint64_t)B * per_thread < n ? begin + (uint64_t)B * per_thread : n;
  uint32_t acc = 0;
  for (uint64_t i = begin + threadIdx.x; i < end; i += B) acc += fold(a[i]);
  if (acc == kSentinel) sink[0] = acc;
}

// Average launch time over `reps` back-to-back launches between two events (after two
// warm-up launches): the same method bench.py uses for the scan, so sustained clocks
// and HBM state are alike.
template <typename Launch>
double back_to_back_ms(Launch launch, int reps) {
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  launch();
  launch();
  CHECK(hipGetLastError());
  CHECK(hipDeviceSynchronize());
  CHECK(hipEventRecord(e0));
  for (int r = 0; r < reps; r++) launch();
  CHECK(hipEventRecord(e1));
  CHECK(hipEventSynchronize(e1));
  float ms = 0;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  return ms / reps;
}

// One wide-layout pattern: columns of W... bytes per TILE-row tile carved one after another
// from `base`, B threads x rows_per_thread rows per block.  Returns ms per pass.
template <int B, int G, int TILE, bool NT, int... W>
double wide_pass(char* base, uint64_t rows, uint32_t rows_per_thread, uint32_t* sink, int reps) {
  using T = WideTile<NT, W...>;
  const int widths[] = {W...};
  WideColumns c{};
  const uint64_t tiles = rows / TILE;
  uint64_t off = 0;
  for (int i = 0; i < T::kColumns; i++) {
    c.col[i] = base + off;
    off += (tiles * (uint64_t)widths[i] + 255) / 256 * 256;
  }
  const uint32_t tpt = rows_per_thread / TILE;
  const uint64_t blocks = (tiles + (uint64_t)B * tpt - 1) / ((uint64_t)B * tpt);
  return back_to_back_ms([&] { read_wide<B, G, T><<<dim3((uint32_t)blocks), dim3(B)>>>(c, tiles, tpt, sink); },
                         reps);
}

}  // namespace

int main(int argc, char** argv) {
  const uint64_t rows = argc > 1 ? std::strtoull(argv[1], nullptr, 10) / 16 * 16 : 600000000ull;
  if (rows < 16 * 4096) {
    std::fprintf(stderr, "hbm_probe: at least 65536 rows\n");
    return 2;
  }
  const uint64_t tiles = rows / 16;
  const uint64_t sizes[7] = {rows * 2, rows, rows * 4, rows, rows, rows / 4, rows / 8};
  uint64_t offs[7], total = 0;
  for (int i = 0; i < 7; i++) {
    offs[i] = total;
    total += (sizes[i] + 255) / 256 * 256;
  }
  char* base = nullptr;
  uint32_t* sink = nullptr;
  CHECK(hipMalloc(&base, total));
  CHECK(hipMalloc(&sink, 4));
  CHECK(hipMemset(base, 0x5A, total));
  Columns c;
  c.ship = reinterpret_cast<const uint4*>(base + offs[0]);
  c.quantity = reinterpret_cast<const uint4*>(base + offs[1]);
  c.price = reinterpret_cast<const uint4*>(base + offs[2]);
  c.discount = reinterpret_cast<const uint4*>(base + offs[3]);
  c.tax = reinterpret_cast<const uint4*>(base + offs[4]);
  c.return_flag = reinterpret_cast<const uint32_t*>(base + offs[5]);
  c.line_status = reinterpret_cast<const uint16_t*>(base + offs[6]);
  const uint64_t bytes = rows * 75 / 8;  // what the column reads touch
  const int reps = 20;

  std::printf("{\"rows\": %llu, \"bytes\": %llu, \"columns\": [", (unsigned long long)rows,
              (unsigned long long)bytes);
  double best_cols = 1e30, best_flat = 1e30;
  const int shapes[7][2] = {{64, 1024}, {64, 4096}, {128, 2048}, {256, 256}, {256, 512}, {256, 1024}, {256, 2048}};
  for (int s = 0; s < 7; s++) {
    const int b = shapes[s][0];
    const uint32_t tpt = shapes[s][1] / 16;  // tiles per thread
    const uint64_t blocks = (tiles + (uint64_t)b * tpt - 1) / ((uint64_t)b * tpt);
    double ms;
    if (b == 64)
      ms = back_to_back_ms([&] { read_columns<64><<<dim3((uint32_t)blocks), dim3(64)>>>(c, tiles, tpt, sink); }, reps);
    else if (b == 128)
      ms = back_to_back_ms([&] { read_columns<128><<<dim3((uint32_t)blocks), dim3(128)>>>(c, tiles, tpt, sink); }, reps);
    else
      ms = back_to_back_ms([&] { read_columns<256><<<dim3((uint32_t)blocks), dim3(256)>>>(c, tiles, tpt, sink); }, reps);
    best_cols = std::min(best_cols, ms);
    std::printf("%s{\"threads_per_block\": %d, \"tuples_per_thread\": %d, \"ms\": %.4f, \"gbps\": %.1f}", s ? ", " : "",
                b, shapes[s][1], ms, bytes / ms / 1e6);
  }
  std::printf("], \"columns_streaming\": [");
  const uint64_t tiles4 = rows / 4;
  double best_stream = 1e30;
  const int stream_shapes[3][2] = {{256, 256}, {256, 1024}, {64, 1024}};
  for (int s = 0; s < 3; s++) {
    const int b = stream_shapes[s][0];
    const uint32_t tpt = stream_shapes[s][1] / 4;  // 4-row tiles per thread
    const uint64_t blocks = (tiles4 + (uint64_t)b * tpt - 1) / ((uint64_t)b * tpt);
    double ms;
    if (b == 64)
      ms = back_to_back_ms([&] { read_columns_nt<64><<<dim3((uint32_t)blocks), dim3(64)>>>(c, tiles4, tpt, sink); }, reps);
    else
      ms = back_to_back_ms([&] { read_columns_nt<256><<<dim3((uint32_t)blocks), dim3(256)>>>(c, tiles4, tpt, sink); }, reps);
    best_stream = std::min(best_stream, ms);
    std::printf("%s{\"threads_per_block\": %d, \"tuples_per_thread\": %d, \"ms\": %.4f, \"gbps\": %.1f}", s ? ", " : "",
                b, stream_shapes[s][1], ms, bytes / ms / 1e6);
  }
  // the batched kernel's tuned structure: 8 blocks per compute unit x 256 lanes, entries of
  // 192 rows per lane taken with a fixed stride
  int cus = 0;
  CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
  const uint32_t res_blocks = 8u * (uint32_t)cus, res_tiles = 192 / 4;
  const double res_ms = back_to_back_ms([&] {
    read_columns_resident<256><<<dim3(res_blocks), dim3(256)>>>(c, tiles4, res_tiles, sink);
  }, reps);
  std::printf("], \"resident_streaming\": {\"blocks\": %u, \"threads_per_block\": 256, \"rows_per_lane\": 192, "
              "\"ms\": %.4f, \"gbps\": %.1f}", res_blocks, res_ms, bytes / res_ms / 1e6);
  std::printf(", \"flat\": [");
  const uint64_t n16 = bytes / 16;
  const int flat[4][2] = {{64, 64}, {256, 64}, {256, 256}, {512, 64}};
  for (int s = 0; s < 4; s++) {
    const int b = flat[s][0];
    const uint32_t pt = flat[s][1];
    const uint64_t blocks = (n16 + (uint64_t)b * pt - 1) / ((uint64_t)b * pt);
    const uint4* a = reinterpret_cast<const uint4*>(base);
    double ms;
    if (b == 64)
      ms = back_to_back_ms([&] { read_flat<64><<<dim3((uint32_t)blocks), dim3(64)>>>(a, n16, pt, sink); }, reps);
    else if (b == 256)
      ms = back_to_back_ms([&] { read_flat<256><<<dim3((uint32_t)blocks), dim3(256)>>>(a, n16, pt, sink); }, reps);
    else
      ms = back_to_back_ms([&] { read_flat<512><<<dim3((uint32_t)blocks), dim3(512)>>>(a, n16, pt, sink); }, reps);
    best_flat = std::min(best_flat, ms);
    std::printf("%s{\"threads_per_block\": %d, \"per_thread\": %u, \"ms\": %.4f, \"gbps\": %.1f}", s ? ", " : "", b, pt,
                ms, n16 * 16 / ms / 1e6);
  }
  best_cols = std::min(best_cols, best_stream);
  CHECK(hipFree(base));

  // The wide layouts over the same rows: uncompressed (20 B/row: int32 ship, int32 qty,
  // int64 price, int8 disc, tax, char rf, ls) and reference (38 B/row: int32 ship, int64
  // qty, price, disc, tax, char rf, ls).
  char* wide = nullptr;
  CHECK(hipMalloc(&wide, rows * 38 + 8 * 256));
  CHECK(hipMemset(wide, 0x5A, rows * 38 + 8 * 256));
  const double unc_bytes = (double)rows * 20, ref_bytes = (double)rows * 38;
  const double u8 = wide_pass<256, 1, 8, false, 32, 32, 64, 8, 8, 8, 8>(wide, rows, 256, sink, reps);
  const double u2 = wide_pass<256, 4, 2, true, 8, 8, 16, 2, 2, 2, 2>(wide, rows, 256, sink, reps);
  const double r4 = wide_pass<256, 1, 4, false, 16, 32, 32, 32, 32, 4, 4>(wide, rows, 256, sink, reps);
  const double r2 = wide_pass<256, 4, 2, true, 8, 16, 16, 16, 16, 2, 2>(wide, rows, 256, sink, reps);
  CHECK(hipFree(wide));

  // Stream copy of 2 GiB into another 2 GiB: bytes read plus bytes written per second.
  const uint64_t copy_bytes = 2ull << 30, n16c = copy_bytes / 16;
  uint4 *csrc = nullptr, *cdst = nullptr;
  CHECK(hipMalloc(&csrc, copy_bytes));
  CHECK(hipMalloc(&cdst, copy_bytes));
  CHECK(hipMemset(csrc, 0x5A, copy_bytes));
  const uint32_t cpt = 64;
  const uint64_t cblocks = (n16c + 256ull * cpt - 1) / (256ull * cpt);
  const double copy_ms =
      back_to_back_ms([&] { copy_flat<256><<<dim3((uint32_t)cblocks), dim3(256)>>>(csrc, cdst, n16c, cpt); }, reps);
  CHECK(hipFree(csrc));
  CHECK(hipFree(cdst));
  const char* fmt = "{\"tile_rows\": %d, \"nontemporal\": %s, \"threads_per_block\": 256, \"tuples_per_thread\": 256, "
                    "\"ms\": %.4f, \"gbps\": %.1f}";
  std::printf("], \"uncompressed\": [");
  std::printf(fmt, 8, "false", u8, unc_bytes / u8 / 1e6);
  std::printf(", ");
  std::printf(fmt, 2, "true", u2, unc_bytes / u2 / 1e6);
  std::printf("], \"reference\": [");
  std::printf(fmt, 4, "false", r4, ref_bytes / r4 / 1e6);
  std::printf(", ");
  std::printf(fmt, 2, "true", r2, ref_bytes / r2 / 1e6);
  std::printf("], \"copy\": {\"bytes\": %llu, \"ms\": %.4f, \"gbps\": %.1f}", (unsigned long long)copy_bytes,
              copy_ms, 2.0 * copy_bytes / copy_ms / 1e6);
  std::printf(", \"resident_gbps\": %.1f", bytes / res_ms / 1e6);
  std::printf(", \"best_columns_gbps\": %.1f, \"best_streaming_gbps\": %.1f, \"best_flat_gbps\": %.1f, "
              "\"best_uncompressed_gbps\": %.1f, \"best_reference_gbps\": %.1f, \"copy_gbps\": %.1f}\n",
              bytes / best_cols / 1e6, bytes / best_stream / 1e6, n16 * 16 / best_flat / 1e6,
              unc_bytes / std::min(u8, u2) / 1e6, ref_bytes / std::min(r4, r2) / 1e6, 2.0 * copy_bytes / copy_ms / 1e6);
  CHECK(hipFree(sink));
  return 0;
}
