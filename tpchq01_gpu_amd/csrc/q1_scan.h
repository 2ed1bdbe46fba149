// gfx950 scan kernels of the TPC-H Q1 scan -> filter -> group-aggregate path: device code
// shared by the translation units that instantiate it, one per column layout
// (q1_scan_<layout>.hip), so the layouts compile in parallel.
//
//
// Work decomposition (all layouts, all placements):
//   * rows are cut into lane tiles of TILE consecutive rows (16 for the compressed layouts,
//     8 for the uncompressed one, 4 for the reference layout); the shared placement reads
//     the compressed layouts in 4-row tiles instead, four in flight per lane, so that every
//     load instruction of a wavefront reads whole cache lines and the loads can be
//     non-temporal;
//   * a block owns tiles_per_thread * blockDim consecutive tiles; in iteration k thread i
//     takes tile base + k*blockDim + i, so each load instruction of a wavefront covers one
//     contiguous stretch of every column (coalesced 16-byte loads);
//   * only the first and the last tile of a range can be partial: those take a guarded
//     per-row path, every other tile reads whole vectors;
//   * each placement keeps its partial aggregation table in its own memory space
//     (registers, scratch, LDS or global memory); at the end of the block the per-lane
//     partials are reduced across the wavefront with ds_bpermute butterflies, across the
//     block through LDS in 128-bit arithmetic, and merged into the destination table with
//     one set of 64-bit atomics per block (128-bit sums as carry-propagating word pairs).
//
// Arithmetic (cpu/kernel_naive.hpp:14-25): disc_1 = 100 - disc, disc_price = disc_1*price,
// charge = disc_price*(100 + tax), all exact.  Inside the TPC-H value domain (discount and
// tax <= 15, quantity <= 63 units, price < 2^24; measured per table) the "fast" kernels use
// 24-bit multiplies and narrow per-tile accumulators:
//   small    = 1 | disc << 5 | qty << 13       (count, discount, quantity in one dword)
//   dprod    = disc * price                     (then sum_disc_price = 100*sum_price - sum_dprod)
//   charge   = (100-disc)*(100+tax) * price     (v_mul_u32_u24 + v_mul_hi_u32_u24)
// Outside the domain the "general" kernel computes in wrapping int64 with 128-bit
// signed sums, which is the reference CPU kernels' arithmetic for any input.
#pragma once

#include <hip/hip_runtime.h>

#include "q1_common.h"
#include "q1_kernels.h"

namespace q1 {
using FastFn = void (*)(ScanArgs, uint32_t, uint64_t*);
using GenFn = void (*)(ScanArgs, uint32_t);

namespace {

constexpr int kMaxBlock = 256;
constexpr int kBatchBlock = (int)kBatchThreads;  // the batched kernel's block size
constexpr uint32_t kNoGroup = 7;

// ------------------------------------------------------------------------------------
// Small helpers
// ------------------------------------------------------------------------------------
// The thread's index in its block (threadIdx.x), hidden from the optimizer by an empty asm.
// Every use reads the index afresh, so values derived from it (lane offsets into the
// columns and the LDS table) are recomputed where they are used instead of being hoisted
// out of the row and entry loops and kept live across them.  That costs a few integer
// operations per tile and saves 8 to 40 vector registers per kernel.  Where that adds a
// wavefront per SIMD it pays only for the batched kernel with in-register sums (179 to 161
// registers, 2 to 3 wavefronts): the 256-request batch takes 1.28 ms instead of 1.67.  The
// scans (compressed LDS table 122 to 106, uncompressed 132 to 112, 3 to 4 wavefronts) and
// the LDS-table batched kernel (148 to 109) measured the same as before, within 0.5 %.
__device__ __forceinline__ uint32_t thread_index() {
  uint32_t t = __builtin_amdgcn_workitem_id_x();
  __asm__ volatile("" : "+v"(t));
  return t;
}
__device__ __forceinline__ uint32_t comp(const uint4& v, int i) {
  return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}
__device__ __forceinline__ void comp_or(uint4& v, int i, uint32_t bits) {
  if (i == 0) v.x |= bits; else if (i == 1) v.y |= bits; else if (i == 2) v.z |= bits; else v.w |= bits;
}
__device__ __forceinline__ void comp_set(uint4& v, int i, uint32_t bits) {
  if (i == 0) v.x = bits; else if (i == 1) v.y = bits; else if (i == 2) v.z = bits; else v.w = bits;
}
__device__ __forceinline__ uint32_t comp(const uint2& v, int i) { return i == 0 ? v.x : v.y; }
__device__ __forceinline__ void comp_set(uint2& v, int i, uint32_t bits) {
  if (i == 0) v.x = bits; else v.y = bits;
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Rows of the tile that lie inside [lo, hi): bit j set for row tile*TILE + j.
template <int TILE>
__device__ __forceinline__ uint32_t valid_mask(uint64_t tile, uint64_t lo, uint64_t hi) {
  const uint64_t r0 = tile * TILE;
  uint32_t m = 0;
#pragma unroll
  for (int j = 0; j < TILE; j++) m |= (uint32_t)(r0 + j >= lo && r0 + j < hi) << j;
  return m;
}

// One row of the fast path: values already narrowed.  gid is the raw 3-bit slot code
// (6 and 7 only arise from the unused return-flag code 3 and never count).
struct FastRow {
  bool pass;     // inside the row range and ship_date <= cutoff
  uint32_t gid;
  uint32_t qty, disc, tax, price;
  __device__ __forceinline__ bool counts() const { return pass && gid < kNumGroups; }
};

// One row of the general path.
struct WideRow {
  uint32_t gid;
  int64_t qty, price, disc, tax;
};

// ------------------------------------------------------------------------------------
// Tile loaders, one per layout.  load() reads a whole tile with vector loads; load_edge()
// reads only the rows inside [lo, hi) (guarded, per row) into the same vector form.
// Accessors extract row j (j is a compile-time constant after unrolling).
// ------------------------------------------------------------------------------------

// Non-temporal (streaming) vector loads.  A column is read once, so its lines need not
// stay in the caches; on gfx950 this raises the rate at which HBM feeds a streaming kernel
// (measured: 6.2 -> 7.1 TB/s for these columns) when every load instruction of a
// wavefront covers whole cache lines, as the 4-row tiles below do.
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint4 nt_load(const uint4* p) {
  const u32x4_t v = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(p));
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint2 nt_load(const uint2* p) {
  const u32x2_t v = __builtin_nontemporal_load(reinterpret_cast<const u32x2_t*>(p));
  return make_uint2(v.x, v.y);
}
__device__ __forceinline__ uint32_t nt_load(const uint32_t* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ uint32_t nt_load(const uint16_t* p) { return __builtin_nontemporal_load(p); }

// Q1_LAYOUT_COMPRESSED / Q1_LAYOUT_COMPRESSED_PUSHDOWN: 16 rows per tile.
template <bool PUSHDOWN>
struct CompressedTile {
  static constexpr int TILE = 16;
  static constexpr int GROUP = 1;
  uint4 sd[2];      // 16 x uint16 ship dates (unused for pushdown)
  uint4 q, d, x;    // 16 x uint8 quantity, discount, tax
  uint4 p[4];       // 16 x uint32 extended price
  uint32_t rf;      // 16 x 2-bit return flag codes
  uint32_t ls;      // 16 x 1-bit line status codes (low half-word)
  uint32_t flt;     // 16 x filter bits (pushdown)

  __device__ __forceinline__ void load(const q1_columns& c, uint64_t t) {
    if (!PUSHDOWN) {
      const uint4* s = reinterpret_cast<const uint4*>(c.ship_date);
      sd[0] = s[2 * t];
      sd[1] = s[2 * t + 1];
    } else {
      flt = reinterpret_cast<const uint16_t*>(c.filter)[t];
    }
    q = reinterpret_cast<const uint4*>(c.quantity)[t];
    d = reinterpret_cast<const uint4*>(c.discount)[t];
    x = reinterpret_cast<const uint4*>(c.tax)[t];
    const uint4* pp = reinterpret_cast<const uint4*>(c.extended_price);
#pragma unroll
    for (int k = 0; k < 4; k++) p[k] = pp[4 * t + k];
    rf = reinterpret_cast<const uint32_t*>(c.return_flag)[t];
    ls = reinterpret_cast<const uint16_t*>(c.line_status)[t];
  }
  __device__ __forceinline__ void load_edge(const q1_columns& c, uint64_t t, uint32_t valid) {
    sd[0] = sd[1] = q = d = x = p[0] = p[1] = p[2] = p[3] = make_uint4(0, 0, 0, 0);
    rf = ls = flt = 0;
    // The packed words of a tile hold rows of the tile only, so reading them for a tile
    // that has at least one valid row stays inside the column.
    if (valid) {
      rf = reinterpret_cast<const uint32_t*>(c.return_flag)[t];
      ls = reinterpret_cast<const uint16_t*>(c.line_status)[t];
      if (PUSHDOWN) flt = reinterpret_cast<const uint16_t*>(c.filter)[t];
    }
#pragma unroll
    for (int j = 0; j < TILE; j++) {
      if (valid >> j & 1) {
        const uint64_t r = t * TILE + j;
        if (!PUSHDOWN) {
          const uint32_t s = reinterpret_cast<const uint16_t*>(c.ship_date)[r];
          comp_or(sd[j >> 3], (j >> 1) & 3, s << (16 * (j & 1)));
        }
        comp_or(q, j >> 2, (uint32_t)reinterpret_cast<const uint8_t*>(c.quantity)[r] << (8 * (j & 3)));
        comp_or(d, j >> 2, (uint32_t)reinterpret_cast<const uint8_t*>(c.discount)[r] << (8 * (j & 3)));
        comp_or(x, j >> 2, (uint32_t)reinterpret_cast<const uint8_t*>(c.tax)[r] << (8 * (j & 3)));
        comp_set(p[j >> 2], j & 3, reinterpret_cast<const uint32_t*>(c.extended_price)[r]);
      }
    }
  }
  // Every row of the tile inside the TPC-H value domain (discount, tax <= 15, quantity
  // <= 63 units, price < 2^24); rows zeroed by load_edge pass.
  __device__ __forceinline__ bool in_domain() const {
    const uint32_t dd = d.x | d.y | d.z | d.w, xx = x.x | x.y | x.z | x.w, qq = q.x | q.y | q.z | q.w;
    uint32_t pp = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) pp |= p[k].x | p[k].y | p[k].z | p[k].w;
    return ((dd | xx) & 0xF0F0F0F0u) == 0 && (qq & 0xC0C0C0C0u) == 0 && (pp & 0xFF000000u) == 0;
  }
  // `cut` is the cutoff relative to the frame of reference (may be < 0 or > 65535).
  __device__ __forceinline__ bool pass(int j, int32_t cut) const {
    if (PUSHDOWN) return (flt >> j) & 1;
    const uint32_t s = (comp(sd[j >> 3], (j >> 1) & 3) >> (16 * (j & 1))) & 0xFFFFu;
    return (int32_t)s <= cut;
  }
  __device__ __forceinline__ uint32_t gid(int j) const {
    return __builtin_amdgcn_ubfe(rf, 2 * j, 2) << 1 | __builtin_amdgcn_ubfe(ls, j, 1);
  }
  __device__ __forceinline__ uint32_t byte(const uint4& v, int j) const {
    return (comp(v, j >> 2) >> (8 * (j & 3))) & 0xFFu;
  }
  // Rows carrying the unused return-flag code 3 (slot 6 or 7): bit 2j set for row j.
  __device__ __forceinline__ uint32_t bad_codes() const { return rf & (rf >> 1) & 0x55555555u; }
  // Turns those rows into rows that fail the filter (ship date 0xFFFF, above every
  // cutoff below 65535; pushdown: filter bit cleared) with code 0, so every row's slot is
  // below 6.
  __device__ void drop_bad_codes() {
    const uint32_t inv = bad_codes();
    rf &= ~(inv * 3u);
#pragma unroll
    for (int j = 0; j < TILE; j++) {
      if (inv >> (2 * j) & 1) {
        if (PUSHDOWN)
          flt &= ~(1u << j);
        else
          comp_or(sd[j >> 3], (j >> 1) & 3, 0xFFFFu << (16 * (j & 1)));
      }
    }
  }
  template <int U>
  __device__ __forceinline__ static void load_group(CompressedTile (&g)[U], const q1_columns& c, uint64_t t0,
                                                    uint64_t stride) {
#pragma unroll
    for (int m = 0; m < U; m++) g[m].load(c, t0 + (uint64_t)m * stride);
  }
  template <bool EDGE>
  __device__ __forceinline__ FastRow fast(int j, int32_t cut, uint32_t valid) const {
    FastRow r;
    r.pass = EDGE ? ((valid >> j & 1) && pass(j, cut)) : pass(j, cut);
    r.gid = gid(j);
    r.qty = byte(q, j);
    r.disc = byte(d, j);
    r.tax = byte(x, j);
    r.price = comp(p[j >> 2], j & 3);
    return r;
  }
  __device__ __forceinline__ WideRow wide(int j, int32_t cut, uint32_t valid) const {
    WideRow r;
    const bool ok = (valid >> j & 1) && pass(j, cut);
    r.gid = ok ? gid(j) : kNoGroup;
    r.qty = byte(q, j);
    r.disc = byte(d, j);
    r.tax = byte(x, j);
    r.price = comp(p[j >> 2], j & 3);
    return r;
  }
};

// Streaming variant for the compressed layouts (used with the shared placement): 4 rows
// per tile, so that each load
// instruction of a wavefront reads one contiguous, line-aligned stretch of its column:
// 64 lanes x 8 bytes of ship dates, x 4 bytes of quantity / discount / tax, x 16 bytes of
// prices.  The 4 rows' return flags are one byte of the packed column (2 bits per row)
// and their line statuses / filter bits one nibble (1 bit per row).
template <bool PUSHDOWN>
struct CompressedStreamTile {
  static constexpr int TILE = 4;
  static constexpr int GROUP = 4;  // tiles loaded together per lane
  uint2 sd;         // 4 x uint16 ship dates (unused for pushdown)
  uint32_t q, d, x; // 4 x uint8 quantity, discount, tax
  uint4 p;          // 4 x uint32 extended price
  uint32_t rf;      // 4 x 2-bit return flag codes
  uint32_t ls;      // 4 x 1-bit line status codes
  uint32_t flt;     // 4 x filter bits (pushdown)

  __device__ __forceinline__ static uint32_t nibble(const void* col, uint64_t t) {
    return (reinterpret_cast<const uint8_t*>(col)[t >> 1] >> (4 * (t & 1))) & 0xFu;
  }
  __device__ __forceinline__ void load(const q1_columns& c, uint64_t t) {
    if (!PUSHDOWN)
      sd = nt_load(reinterpret_cast<const uint2*>(c.ship_date) + t);
    else
      flt = nibble(c.filter, t);
    q = nt_load(reinterpret_cast<const uint32_t*>(c.quantity) + t);
    d = nt_load(reinterpret_cast<const uint32_t*>(c.discount) + t);
    x = nt_load(reinterpret_cast<const uint32_t*>(c.tax) + t);
    p = nt_load(reinterpret_cast<const uint4*>(c.extended_price) + t);
    rf = reinterpret_cast<const uint8_t*>(c.return_flag)[t];
    ls = nibble(c.line_status, t);
  }
  __device__ __forceinline__ void load_edge(const q1_columns& c, uint64_t t, uint32_t valid) {
    sd = make_uint2(0, 0);
    q = d = x = 0;
    p = make_uint4(0, 0, 0, 0);
    rf = ls = flt = 0;
    // The packed byte / nibble of a tile holds rows of the tile only, so reading it for a
    // tile that has at least one valid row stays inside the column.
    if (valid) {
      rf = reinterpret_cast<const uint8_t*>(c.return_flag)[t];
      ls = nibble(c.line_status, t);
      if (PUSHDOWN) flt = nibble(c.filter, t);
    }
#pragma unroll
    for (int j = 0; j < TILE; j++) {
      if (valid >> j & 1) {
        const uint64_t r = t * TILE + j;
        if (!PUSHDOWN) {
          const uint32_t s = reinterpret_cast<const uint16_t*>(c.ship_date)[r];
          if (j < 2) sd.x |= s << (16 * (j & 1)); else sd.y |= s << (16 * (j & 1));
        }
        q |= (uint32_t)reinterpret_cast<const uint8_t*>(c.quantity)[r] << (8 * j);
        d |= (uint32_t)reinterpret_cast<const uint8_t*>(c.discount)[r] << (8 * j);
        x |= (uint32_t)reinterpret_cast<const uint8_t*>(c.tax)[r] << (8 * j);
        comp_set(p, j, reinterpret_cast<const uint32_t*>(c.extended_price)[r]);
      }
    }
  }
  // Every row of the tile inside the TPC-H value domain (discount, tax <= 15, quantity
  // <= 63 units, price < 2^24); rows zeroed by load_edge pass.
  __device__ __forceinline__ bool in_domain() const {
    return ((d | x) & 0xF0F0F0F0u) == 0 && (q & 0xC0C0C0C0u) == 0 &&
           ((p.x | p.y | p.z | p.w) & 0xFF000000u) == 0;
  }
  // `cut` is the cutoff relative to the frame of reference (may be < 0 or > 65535).
  __device__ __forceinline__ bool pass(int j, int32_t cut) const {
    if (PUSHDOWN) return (flt >> j) & 1;
    const uint32_t s = ((j < 2 ? sd.x : sd.y) >> (16 * (j & 1))) & 0xFFFFu;
    return (int32_t)s <= cut;
  }
  __device__ __forceinline__ uint32_t gid(int j) const {
    return __builtin_amdgcn_ubfe(rf, 2 * j, 2) << 1 | __builtin_amdgcn_ubfe(ls, j, 1);
  }
  __device__ __forceinline__ static uint32_t byte(uint32_t v, int j) { return (v >> (8 * j)) & 0xFFu; }
  // Rows carrying the unused return-flag code 3 (slot 6 or 7): bit 2j set for row j.
  __device__ __forceinline__ uint32_t bad_codes() const { return rf & (rf >> 1) & 0x55u; }
  // Turns those rows into rows that fail the filter (ship date 0xFFFF, above every
  // cutoff below 65535; pushdown: filter bit cleared) with code 0, so every row's slot is
  // below 6.
  __device__ void drop_bad_codes() {
    const uint32_t inv = bad_codes();
    rf &= ~(inv * 3u);
    if (PUSHDOWN) {
      flt &= ~((inv & 1u) | (inv >> 1 & 2u) | (inv >> 2 & 4u) | (inv >> 3 & 8u));
    } else {
      sd.x |= (inv & 1u ? 0xFFFFu : 0u) | (inv & 4u ? 0xFFFF0000u : 0u);
      sd.y |= (inv & 0x10u ? 0xFFFFu : 0u) | (inv & 0x40u ? 0xFFFF0000u : 0u);
    }
  }
  // U tiles a block-stride apart (stride even): their packed line-status / filter nibbles
  // sit stride / 2 bytes apart from one base and at the same half of their byte.
  template <int U>
  __device__ __forceinline__ static void load_group(CompressedStreamTile (&g)[U], const q1_columns& c, uint64_t t0,
                                                    uint64_t stride) {
    const uint8_t* lsb = reinterpret_cast<const uint8_t*>(c.line_status) + (t0 >> 1);
    const uint32_t sh = 4u * (uint32_t)(t0 & 1);
#pragma unroll
    for (int m = 0; m < U; m++) {
      const uint64_t t = t0 + (uint64_t)m * stride;
      if (!PUSHDOWN)
        g[m].sd = nt_load(reinterpret_cast<const uint2*>(c.ship_date) + t);
      else
        g[m].flt = (reinterpret_cast<const uint8_t*>(c.filter)[(t0 >> 1) + (uint64_t)m * (stride >> 1)] >> sh) & 0xFu;
      g[m].q = nt_load(reinterpret_cast<const uint32_t*>(c.quantity) + t);
      g[m].d = nt_load(reinterpret_cast<const uint32_t*>(c.discount) + t);
      g[m].x = nt_load(reinterpret_cast<const uint32_t*>(c.tax) + t);
      g[m].p = nt_load(reinterpret_cast<const uint4*>(c.extended_price) + t);
      g[m].rf = reinterpret_cast<const uint8_t*>(c.return_flag)[t];
      g[m].ls = (lsb[(uint64_t)m * (stride >> 1)] >> sh) & 0xFu;
    }
  }
  template <bool EDGE>
  __device__ __forceinline__ FastRow fast(int j, int32_t cut, uint32_t valid) const {
    FastRow r;
    r.pass = EDGE ? ((valid >> j & 1) && pass(j, cut)) : pass(j, cut);
    r.gid = gid(j);
    r.qty = byte(q, j);
    r.disc = byte(d, j);
    r.tax = byte(x, j);
    r.price = comp(p, j);
    return r;
  }
  __device__ __forceinline__ WideRow wide(int j, int32_t cut, uint32_t valid) const {
    WideRow r;
    const bool ok = (valid >> j & 1) && pass(j, cut);
    r.gid = ok ? gid(j) : kNoGroup;
    r.qty = byte(q, j);
    r.disc = byte(d, j);
    r.tax = byte(x, j);
    r.price = comp(p, j);
    return r;
  }
};

// Q1_LAYOUT_UNCOMPRESSED: int32 ship, int32 qty, int64 price, int8 disc/tax, char rf/ls.
struct UncompressedTile {
  static constexpr int TILE = 8;
  static constexpr int GROUP = 1;
  uint4 s[2], q[2], p[4];
  uint2 d, x, rf, ls;

  __device__ __forceinline__ void load(const q1_columns& c, uint64_t t) {
    const uint4* ss = reinterpret_cast<const uint4*>(c.ship_date);
    const uint4* qq = reinterpret_cast<const uint4*>(c.quantity);
    const uint4* pp = reinterpret_cast<const uint4*>(c.extended_price);
    s[0] = ss[2 * t]; s[1] = ss[2 * t + 1];
    q[0] = qq[2 * t]; q[1] = qq[2 * t + 1];
#pragma unroll
    for (int k = 0; k < 4; k++) p[k] = pp[4 * t + k];
    d = reinterpret_cast<const uint2*>(c.discount)[t];
    x = reinterpret_cast<const uint2*>(c.tax)[t];
    rf = reinterpret_cast<const uint2*>(c.return_flag)[t];
    ls = reinterpret_cast<const uint2*>(c.line_status)[t];
  }
  __device__ __forceinline__ void load_edge(const q1_columns& c, uint64_t t, uint32_t valid) {
    s[0] = s[1] = q[0] = q[1] = p[0] = p[1] = p[2] = p[3] = make_uint4(0, 0, 0, 0);
    d = x = rf = ls = make_uint2(0, 0);
#pragma unroll
    for (int j = 0; j < TILE; j++) {
      if (valid >> j & 1) {
        const uint64_t r = t * TILE + j;
        comp_set(s[j >> 2], j & 3, (uint32_t)reinterpret_cast<const int32_t*>(c.ship_date)[r]);
        comp_set(q[j >> 2], j & 3, (uint32_t)reinterpret_cast<const int32_t*>(c.quantity)[r]);
        const uint64_t pr = (uint64_t)reinterpret_cast<const int64_t*>(c.extended_price)[r];
        comp_set(p[j >> 1], 2 * (j & 1), (uint32_t)pr);
        comp_set(p[j >> 1], 2 * (j & 1) + 1, (uint32_t)(pr >> 32));
        const int sh = 8 * (j & 3);
        uint32_t* dd = (j < 4) ? &d.x : &d.y;
        uint32_t* xx = (j < 4) ? &x.x : &x.y;
        uint32_t* rr = (j < 4) ? &rf.x : &rf.y;
        uint32_t* ll = (j < 4) ? &ls.x : &ls.y;
        *dd |= (uint32_t)reinterpret_cast<const uint8_t*>(c.discount)[r] << sh;
        *xx |= (uint32_t)reinterpret_cast<const uint8_t*>(c.tax)[r] << sh;
        *rr |= (uint32_t)reinterpret_cast<const uint8_t*>(c.return_flag)[r] << sh;
        *ll |= (uint32_t)reinterpret_cast<const uint8_t*>(c.line_status)[r] << sh;
      }
    }
  }
  __device__ __forceinline__ static uint32_t byte8(const uint2& v, int j) {
    return ((j < 4 ? v.x : v.y) >> (8 * (j & 3))) & 0xFFu;
  }
  __device__ __forceinline__ uint32_t gid(int j) const {
    return encode_return_flag(byte8(rf, j)) << 1 | encode_line_status(byte8(ls, j));
  }
  __device__ __forceinline__ bool pass(int j, int32_t cut) const {
    return (int32_t)comp(s[j >> 2], j & 3) <= cut;
  }
  // Every row inside the TPC-H value domain (discount, tax in [0, 15], quantity in
  // [0, 6300] hundredths, price in [0, 2^24)); rows zeroed by load_edge pass.
  __device__ __forceinline__ bool in_domain() const {
    uint32_t qmax = 0, phi = 0, plo = 0;
#pragma unroll
    for (int k = 0; k < 2; k++) qmax = max(qmax, max(max(q[k].x, q[k].y), max(q[k].z, q[k].w)));
#pragma unroll
    for (int k = 0; k < 4; k++) {
      plo |= p[k].x | p[k].z;
      phi |= p[k].y | p[k].w;
    }
    return qmax <= 6300u && phi == 0 && (plo & 0xFF000000u) == 0 && ((d.x | d.y | x.x | x.y) & 0xF0F0F0F0u) == 0;
  }
  // Character flags encode to slots 0..5 only.
  __device__ __forceinline__ static constexpr uint32_t bad_codes() { return 0; }
  __device__ __forceinline__ void drop_bad_codes() {}
  template <int U>
  __device__ __forceinline__ static void load_group(UncompressedTile (&g)[U], const q1_columns& c, uint64_t t0,
                                                    uint64_t stride) {
#pragma unroll
    for (int m = 0; m < U; m++) g[m].load(c, t0 + (uint64_t)m * stride);
  }
  template <bool EDGE>
  __device__ __forceinline__ FastRow fast(int j, int32_t cut, uint32_t valid) const {
    FastRow r;
    r.pass = EDGE ? ((valid >> j & 1) && pass(j, cut)) : pass(j, cut);
    r.gid = gid(j);
    r.qty = comp(q[j >> 2], j & 3);
    r.disc = byte8(d, j);
    r.tax = byte8(x, j);
    r.price = comp(p[j >> 1], 2 * (j & 1));
    return r;
  }
  __device__ __forceinline__ WideRow wide(int j, int32_t cut, uint32_t valid) const {
    WideRow r;
    const bool ok = (valid >> j & 1) && pass(j, cut);
    r.gid = ok ? gid(j) : kNoGroup;
    r.qty = (int32_t)comp(q[j >> 2], j & 3);
    r.disc = (int8_t)byte8(d, j);
    r.tax = (int8_t)byte8(x, j);
    r.price = (int64_t)((uint64_t)comp(p[j >> 1], 2 * (j & 1) + 1) << 32 | comp(p[j >> 1], 2 * (j & 1)));
    return r;
  }
};

// Streaming variant for the uncompressed layout (used with the shared placement): 2 rows
// per tile, so that each load instruction of a wavefront reads one contiguous stretch of
// its column with at most 16 bytes per lane: 64 lanes x 8 bytes of ship dates and
// quantities, x 16 bytes of prices, x 2 bytes of discounts, taxes, return flags and line
// statuses.  All loads are non-temporal (each byte is read once).
struct UncompressedStreamTile {
  static constexpr int TILE = 2;
  static constexpr int GROUP = 4;  // tiles loaded together per lane
  uint2 s, q;              // 2 x int32 ship date, quantity (hundredths)
  uint4 p;                 // 2 x int64 price as (lo, hi) pairs
  uint32_t d, x, rf, ls;   // 2 x int8 / char in the low half-word

  __device__ __forceinline__ void load(const q1_columns& c, uint64_t t) {
    s = nt_load(reinterpret_cast<const uint2*>(c.ship_date) + t);
    q = nt_load(reinterpret_cast<const uint2*>(c.quantity) + t);
    p = nt_load(reinterpret_cast<const uint4*>(c.extended_price) + t);
    d = nt_load(reinterpret_cast<const uint16_t*>(c.discount) + t);
    x = nt_load(reinterpret_cast<const uint16_t*>(c.tax) + t);
    rf = nt_load(reinterpret_cast<const uint16_t*>(c.return_flag) + t);
    ls = nt_load(reinterpret_cast<const uint16_t*>(c.line_status) + t);
  }
  __device__ __forceinline__ void load_edge(const q1_columns& c, uint64_t t, uint32_t valid) {
    s = q = make_uint2(0, 0);
    p = make_uint4(0, 0, 0, 0);
    d = x = rf = ls = 0;
#pragma unroll
    for (int j = 0; j < TILE; j++) {
      if (valid >> j & 1) {
        const uint64_t r = t * TILE + j;
        const uint32_t sv = (uint32_t)reinterpret_cast<const int32_t*>(c.ship_date)[r];
        const uint32_t qv = (uint32_t)reinterpret_cast<const int32_t*>(c.quantity)[r];
        const uint64_t pv = (uint64_t)reinterpret_cast<const int64_t*>(c.extended_price)[r];
        comp_set(s, j, sv);
        comp_set(q, j, qv);
        comp_set(p, 2 * j, (uint32_t)pv);
        comp_set(p, 2 * j + 1, (uint32_t)(pv >> 32));
        d |= (uint32_t)reinterpret_cast<const uint8_t*>(c.discount)[r] << (8 * j);
        x |= (uint32_t)reinterpret_cast<const uint8_t*>(c.tax)[r] << (8 * j);
        rf |= (uint32_t)reinterpret_cast<const uint8_t*>(c.return_flag)[r] << (8 * j);
        ls |= (uint32_t)reinterpret_cast<const uint8_t*>(c.line_status)[r] << (8 * j);
      }
    }
  }
  __device__ __forceinline__ static uint32_t byte(uint32_t v, int j) { return (v >> (8 * j)) & 0xFFu; }
  __device__ __forceinline__ uint32_t gid(int j) const {
    return encode_return_flag(byte(rf, j)) << 1 | encode_line_status(byte(ls, j));
  }
  __device__ __forceinline__ bool pass(int j, int32_t cut) const { return (int32_t)comp(s, j) <= cut; }
  // Every row inside the TPC-H value domain (discount, tax in [0, 15], quantity in
  // [0, 6300] hundredths, price in [0, 2^24)); rows zeroed by load_edge pass.
  __device__ __forceinline__ bool in_domain() const {
    return max(q.x, q.y) <= 6300u && (p.y | p.w) == 0 && ((p.x | p.z) & 0xFF000000u) == 0 &&
           ((d | x) & 0xF0F0u) == 0;
  }
  // Character flags encode to slots 0..5 only.
  __device__ __forceinline__ static constexpr uint32_t bad_codes() { return 0; }
  __device__ __forceinline__ void drop_bad_codes() {}
  template <int U>
  __device__ __forceinline__ static void load_group(UncompressedStreamTile (&g)[U], const q1_columns& c, uint64_t t0,
                                                    uint64_t stride) {
#pragma unroll
    for (int m = 0; m < U; m++) g[m].load(c, t0 + (uint64_t)m * stride);
  }
  template <bool EDGE>
  __device__ __forceinline__ FastRow fast(int j, int32_t cut, uint32_t valid) const {
    FastRow r;
    r.pass = EDGE ? ((valid >> j & 1) && pass(j, cut)) : pass(j, cut);
    r.gid = gid(j);
    r.qty = comp(q, j);
    r.disc = byte(d, j);
    r.tax = byte(x, j);
    r.price = comp(p, 2 * j);
    return r;
  }
  __device__ __forceinline__ WideRow wide(int j, int32_t cut, uint32_t valid) const {
    WideRow r;
    const bool ok = (valid >> j & 1) && pass(j, cut);
    r.gid = ok ? gid(j) : kNoGroup;
    r.qty = (int32_t)comp(q, j);
    r.disc = (int8_t)byte(d, j);
    r.tax = (int8_t)byte(x, j);
    r.price = (int64_t)((uint64_t)comp(p, 2 * j + 1) << 32 | comp(p, 2 * j));
    return r;
  }
};

// Q1_LAYOUT_REFERENCE: int32 ship, int64 qty/price/disc/tax, char rf/ls.
struct ReferenceTile {
  static constexpr int TILE = 4;
  static constexpr int GROUP = 1;
  uint4 s, q[2], p[2], d[2], x[2];
  uint32_t rf, ls;

  __device__ __forceinline__ void load(const q1_columns& c, uint64_t t) {
    s = reinterpret_cast<const uint4*>(c.ship_date)[t];
    const uint4* qq = reinterpret_cast<const uint4*>(c.quantity);
    const uint4* pp = reinterpret_cast<const uint4*>(c.extended_price);
    const uint4* dd = reinterpret_cast<const uint4*>(c.discount);
    const uint4* xx = reinterpret_cast<const uint4*>(c.tax);
#pragma unroll
    for (int k = 0; k < 2; k++) {
      q[k] = qq[2 * t + k];
      p[k] = pp[2 * t + k];
      d[k] = dd[2 * t + k];
      x[k] = xx[2 * t + k];
    }
    rf = reinterpret_cast<const uint32_t*>(c.return_flag)[t];
    ls = reinterpret_cast<const uint32_t*>(c.line_status)[t];
  }
  __device__ __forceinline__ static void put64(uint4* v, int j, uint64_t val) {
    comp_set(v[j >> 1], 2 * (j & 1), (uint32_t)val);
    comp_set(v[j >> 1], 2 * (j & 1) + 1, (uint32_t)(val >> 32));
  }
  __device__ __forceinline__ void load_edge(const q1_columns& c, uint64_t t, uint32_t valid) {
    s = q[0] = q[1] = p[0] = p[1] = d[0] = d[1] = x[0] = x[1] = make_uint4(0, 0, 0, 0);
    rf = ls = 0;
#pragma unroll
    for (int j = 0; j < TILE; j++) {
      if (valid >> j & 1) {
        const uint64_t r = t * TILE + j;
        comp_set(s, j, (uint32_t)reinterpret_cast<const int32_t*>(c.ship_date)[r]);
        put64(q, j, (uint64_t)reinterpret_cast<const int64_t*>(c.quantity)[r]);
        put64(p, j, (uint64_t)reinterpret_cast<const int64_t*>(c.extended_price)[r]);
        put64(d, j, (uint64_t)reinterpret_cast<const int64_t*>(c.discount)[r]);
        put64(x, j, (uint64_t)reinterpret_cast<const int64_t*>(c.tax)[r]);
        rf |= (uint32_t)reinterpret_cast<const uint8_t*>(c.return_flag)[r] << (8 * j);
        ls |= (uint32_t)reinterpret_cast<const uint8_t*>(c.line_status)[r] << (8 * j);
      }
    }
  }
  __device__ __forceinline__ static uint32_t lo32(const uint4* v, int j) {
    return comp(v[j >> 1], 2 * (j & 1));
  }
  __device__ __forceinline__ static int64_t get64(const uint4* v, int j) {
    return (int64_t)((uint64_t)comp(v[j >> 1], 2 * (j & 1) + 1) << 32 | comp(v[j >> 1], 2 * (j & 1)));
  }
  __device__ __forceinline__ uint32_t gid(int j) const {
    return encode_return_flag((rf >> (8 * j)) & 0xFFu) << 1 | encode_line_status((ls >> (8 * j)) & 0xFFu);
  }
  __device__ __forceinline__ bool pass(int j, int32_t cut) const { return (int32_t)comp(s, j) <= cut; }
  // Every row inside the TPC-H value domain (int64 discount, tax in [0, 15], quantity in
  // [0, 6300], price in [0, 2^24)); rows zeroed by load_edge pass.
  __device__ __forceinline__ bool in_domain() const {
    uint32_t hi = 0, qmax = 0, plo = 0, dx = 0;
#pragma unroll
    for (int k = 0; k < 2; k++) {
      hi |= q[k].y | q[k].w | p[k].y | p[k].w | d[k].y | d[k].w | x[k].y | x[k].w;
      qmax = max(qmax, max(q[k].x, q[k].z));
      plo |= p[k].x | p[k].z;
      dx |= d[k].x | d[k].z | x[k].x | x[k].z;
    }
    return hi == 0 && qmax <= 6300u && (plo & 0xFF000000u) == 0 && (dx & ~15u) == 0;
  }
  // Character flags encode to slots 0..5 only.
  __device__ __forceinline__ static constexpr uint32_t bad_codes() { return 0; }
  __device__ __forceinline__ void drop_bad_codes() {}
  template <int U>
  __device__ __forceinline__ static void load_group(ReferenceTile (&g)[U], const q1_columns& c, uint64_t t0,
                                                    uint64_t stride) {
#pragma unroll
    for (int m = 0; m < U; m++) g[m].load(c, t0 + (uint64_t)m * stride);
  }
  template <bool EDGE>
  __device__ __forceinline__ FastRow fast(int j, int32_t cut, uint32_t valid) const {
    FastRow r;
    r.pass = EDGE ? ((valid >> j & 1) && pass(j, cut)) : pass(j, cut);
    r.gid = gid(j);
    r.qty = lo32(q, j);
    r.disc = lo32(d, j);
    r.tax = lo32(x, j);
    r.price = lo32(p, j);
    return r;
  }
  __device__ __forceinline__ WideRow wide(int j, int32_t cut, uint32_t valid) const {
    WideRow r;
    const bool ok = (valid >> j & 1) && pass(j, cut);
    r.gid = ok ? gid(j) : kNoGroup;
    r.qty = get64(q, j);
    r.disc = get64(d, j);
    r.tax = get64(x, j);
    r.price = get64(p, j);
    return r;
  }
};

// Streaming variant for the reference layout (used with the shared placement): 2 rows per
// tile, each load instruction of a wavefront one contiguous stretch of its column with at
// most 16 bytes per lane (64 lanes x 8 bytes of ship dates, x 16 bytes of each int64
// column, x 2 bytes of return flags and line statuses), all loads non-temporal.
struct ReferenceStreamTile {
  static constexpr int TILE = 2;
  static constexpr int GROUP = 4;  // tiles loaded together per lane
  uint2 s;                 // 2 x int32 ship date
  uint4 q, p, d, x;        // 2 x int64 as (lo, hi) pairs
  uint32_t rf, ls;         // 2 x char in the low half-word

  __device__ __forceinline__ void load(const q1_columns& c, uint64_t t) {
    s = nt_load(reinterpret_cast<const uint2*>(c.ship_date) + t);
    q = nt_load(reinterpret_cast<const uint4*>(c.quantity) + t);
    p = nt_load(reinterpret_cast<const uint4*>(c.extended_price) + t);
    d = nt_load(reinterpret_cast<const uint4*>(c.discount) + t);
    x = nt_load(reinterpret_cast<const uint4*>(c.tax) + t);
    rf = nt_load(reinterpret_cast<const uint16_t*>(c.return_flag) + t);
    ls = nt_load(reinterpret_cast<const uint16_t*>(c.line_status) + t);
  }
  __device__ __forceinline__ static void put64(uint4& v, int j, uint64_t val) {
    comp_set(v, 2 * j, (uint32_t)val);
    comp_set(v, 2 * j + 1, (uint32_t)(val >> 32));
  }
  __device__ __forceinline__ void load_edge(const q1_columns& c, uint64_t t, uint32_t valid) {
    s = make_uint2(0, 0);
    q = p = d = x = make_uint4(0, 0, 0, 0);
    rf = ls = 0;
#pragma unroll
    for (int j = 0; j < TILE; j++) {
      if (valid >> j & 1) {
        const uint64_t r = t * TILE + j;
        comp_set(s, j, (uint32_t)reinterpret_cast<const int32_t*>(c.ship_date)[r]);
        put64(q, j, (uint64_t)reinterpret_cast<const int64_t*>(c.quantity)[r]);
        put64(p, j, (uint64_t)reinterpret_cast<const int64_t*>(c.extended_price)[r]);
        put64(d, j, (uint64_t)reinterpret_cast<const int64_t*>(c.discount)[r]);
        put64(x, j, (uint64_t)reinterpret_cast<const int64_t*>(c.tax)[r]);
        rf |= (uint32_t)reinterpret_cast<const uint8_t*>(c.return_flag)[r] << (8 * j);
        ls |= (uint32_t)reinterpret_cast<const uint8_t*>(c.line_status)[r] << (8 * j);
      }
    }
  }
  __device__ __forceinline__ static uint32_t lo32(const uint4& v, int j) { return comp(v, 2 * j); }
  __device__ __forceinline__ static int64_t get64(const uint4& v, int j) {
    return (int64_t)((uint64_t)comp(v, 2 * j + 1) << 32 | comp(v, 2 * j));
  }
  __device__ __forceinline__ uint32_t gid(int j) const {
    return encode_return_flag((rf >> (8 * j)) & 0xFFu) << 1 | encode_line_status((ls >> (8 * j)) & 0xFFu);
  }
  __device__ __forceinline__ bool pass(int j, int32_t cut) const { return (int32_t)comp(s, j) <= cut; }
  // Every row inside the TPC-H value domain (int64 discount, tax in [0, 15], quantity in
  // [0, 6300], price in [0, 2^24)); rows zeroed by load_edge pass.
  __device__ __forceinline__ bool in_domain() const {
    const uint32_t hi = q.y | q.w | p.y | p.w | d.y | d.w | x.y | x.w;
    return hi == 0 && max(q.x, q.z) <= 6300u && ((p.x | p.z) & 0xFF000000u) == 0 &&
           ((d.x | d.z | x.x | x.z) & ~15u) == 0;
  }
  // Character flags encode to slots 0..5 only.
  __device__ __forceinline__ static constexpr uint32_t bad_codes() { return 0; }
  __device__ __forceinline__ void drop_bad_codes() {}
  template <int U>
  __device__ __forceinline__ static void load_group(ReferenceStreamTile (&g)[U], const q1_columns& c, uint64_t t0,
                                                    uint64_t stride) {
#pragma unroll
    for (int m = 0; m < U; m++) g[m].load(c, t0 + (uint64_t)m * stride);
  }
  template <bool EDGE>
  __device__ __forceinline__ FastRow fast(int j, int32_t cut, uint32_t valid) const {
    FastRow r;
    r.pass = EDGE ? ((valid >> j & 1) && pass(j, cut)) : pass(j, cut);
    r.gid = gid(j);
    r.qty = lo32(q, j);
    r.disc = lo32(d, j);
    r.tax = lo32(x, j);
    r.price = lo32(p, j);
    return r;
  }
  __device__ __forceinline__ WideRow wide(int j, int32_t cut, uint32_t valid) const {
    WideRow r;
    const bool ok = (valid >> j & 1) && pass(j, cut);
    r.gid = ok ? gid(j) : kNoGroup;
    r.qty = get64(q, j);
    r.disc = get64(d, j);
    r.tax = get64(x, j);
    r.price = get64(p, j);
    return r;
  }
};

template <int L> struct TileOf;
template <> struct TileOf<Q1_LAYOUT_REFERENCE> { using T = ReferenceTile; };
template <> struct TileOf<Q1_LAYOUT_UNCOMPRESSED> { using T = UncompressedTile; };
template <> struct TileOf<Q1_LAYOUT_COMPRESSED> { using T = CompressedTile<false>; };
template <> struct TileOf<Q1_LAYOUT_COMPRESSED_PUSHDOWN> { using T = CompressedTile<true>; };

// The tile a scan_fast<L, P> kernel reads.  The LDS placement (little register state per
// lane) streams small tiles with non-temporal loads, four tiles in flight per lane: 4 rows
// for the compressed layouts, 2 for uncompressed and reference (specialisations below);
// the register placements keep TileOf<L>, whose tile sizes are also the units the host
// counts tiles in (tile_rows).
template <int L, int P> struct ScanTileOf { using T = typename TileOf<L>::T; };

// A group of U lane tiles t0 + m * stride (m < U, at most `count` of them, below
// `end_tile`), all loaded before any of their rows is used, so that a wavefront keeps
// 16 rows' worth of loads per lane in flight whatever the tile size.
template <class Tile, int U>
struct TileGroup {
  Tile tile[U];
  uint32_t valid[U];  // rows of the tile inside [row_begin, row_end)
  bool use[U], edge[U];
  bool plain;         // every tile used and whole: loaded without any branch
  // Straight-line loads of all U tiles, for a group the caller knows to be whole and inside
  // the range for every lane.
  __device__ __forceinline__ void load_plain(const q1_columns& c, uint64_t t0, uint64_t stride) {
    Tile::template load_group<U>(tile, c, t0, stride);
  }
  __device__ __forceinline__ void load(const q1_columns& c, uint64_t t0, uint64_t stride, uint32_t count,
                                       uint64_t end_tile, uint64_t row_begin, uint64_t row_end) {
    constexpr int TILE = Tile::TILE;
    const uint64_t t_last = t0 + (uint64_t)(U - 1) * stride;
    plain = count >= (uint32_t)U && t_last < end_tile && t0 * TILE >= row_begin && (t_last + 1) * TILE <= row_end;
    if (plain) {
      // straight-line loads of all U tiles: every load of the group is in flight at once
#pragma unroll
      for (int m = 0; m < U; m++) {
        tile[m].load(c, t0 + (uint64_t)m * stride);
        use[m] = true;
        edge[m] = false;
        valid[m] = (1u << TILE) - 1;
      }
      return;
    }
#pragma unroll
    for (int m = 0; m < U; m++) {
      const uint64_t t = t0 + (uint64_t)m * stride;
      use[m] = (uint32_t)m < count && t < end_tile;
      edge[m] = t * TILE < row_begin || (t + 1) * TILE > row_end;
      valid[m] = (1u << TILE) - 1;
      if (use[m]) {
        if (!edge[m]) {
          tile[m].load(c, t);
        } else {
          valid[m] = valid_mask<TILE>(t, row_begin, row_end);
          tile[m].load_edge(c, t, valid[m]);
        }
      }
    }
  }
};

// Cutoff in the units of the layout's ship_date column.
template <int L>
__device__ __forceinline__ int32_t column_cutoff(int32_t cutoff_day) {
  if (L == Q1_LAYOUT_COMPRESSED || L == Q1_LAYOUT_COMPRESSED_PUSHDOWN) {
    const int64_t c = (int64_t)cutoff_day - kShipDateFrameOfReference;
    return c < -1 ? -1 : (c > 65535 ? 65535 : (int32_t)c);
  }
  return cutoff_day;
}

// ------------------------------------------------------------------------------------
// Per-lane partial table handed to the block merge: six base sums per group.
//   F_QTY quantity (layout units), F_PRICE price, F_DISC discount, F_CNT count,
//   F_DPROD sum of disc*price, F_CHARGE charge.  All exact in uint64 per lane.
// ------------------------------------------------------------------------------------
enum { F_QTY, F_PRICE, F_DISC, F_CNT, F_DPROD, F_CHARGE, NF };

struct LaneTotals {
  uint64_t v[kNumGroups][NF];
};

__device__ __forceinline__ uint32_t pack_small(const FastRow& r) {
  return 1u | r.disc << 5 | r.qty << 13;
}
__device__ __forceinline__ uint32_t disc_prod(const FastRow& r) { return __umul24(r.disc, r.price); }
__device__ __forceinline__ uint64_t fast_charge(const FastRow& r) {
  const uint32_t f = __umul24(100u - r.disc, 100u + r.tax);  // <= 100*115, fits 24 bits
  const uint64_t prod = (uint64_t)(f & 0xFFFFFFu) * (r.price & 0xFFFFFFu);
  return prod;
}

// DPP move of a 64-bit value (two 32-bit lane permutes); CTRL is a gfx9 DPP control.
template <int CTRL>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
  const uint2 w = __builtin_bit_cast(uint2, v);
  const uint2 r = make_uint2((uint32_t)__builtin_amdgcn_mov_dpp((int)w.x, CTRL, 0xF, 0xF, false),
                             (uint32_t)__builtin_amdgcn_mov_dpp((int)w.y, CTRL, 0xF, 0xF, false));
  return __builtin_bit_cast(uint64_t, r);
}
__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int lane) {
  const uint2 w = __builtin_bit_cast(uint2, v);
  const uint2 r = make_uint2((uint32_t)__builtin_amdgcn_readlane((int)w.x, lane),
                             (uint32_t)__builtin_amdgcn_readlane((int)w.y, lane));
  return __builtin_bit_cast(uint64_t, r);
}

// Sum of a 64-bit value over the 64 lanes of a wavefront (all lanes active); the result
// is wavefront-uniform.  Within each row of 16 lanes: quad butterflies (quad_perm
// [1,0,3,2] and [2,3,0,1]) then rotations by 4 and 8 leave the row total in every lane;
// the four row totals are read into scalar registers and added there.
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
  v += dpp_u64<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_u64<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_u64<0x124>(v);  // row_ror:4
  v += dpp_u64<0x128>(v);  // row_ror:8
  return readlane_u64(v, 0) + readlane_u64(v, 16) + readlane_u64(v, 32) + readlane_u64(v, 48);
}

// Q1_PLACE_IN_REGISTERS: narrow per-tile accumulators in registers; a group's block of
// adds is skipped by the whole wavefront when no lane holds a row of that group.  Price
// and price x discount share one 64-bit word (price in the low half, below 16 * 2^24 =
// 2^28 over a tile; the product in the high half, below 16 * 15 * 2^24 < 2^32), so a row
// costs three adds per group: that word, the packed count|discount|quantity and the charge.
struct AccRegisters {
  static constexpr bool kPackedSmall = false;
  uint32_t ns[kNumGroups];
  uint64_t npd[kNumGroups];
  uint64_t ch[kNumGroups], pr[kNumGroups], dp[kNumGroups];
  uint32_t cnt[kNumGroups], dsc[kNumGroups], qty[kNumGroups];

  __device__ __forceinline__ void init(char*) {
#pragma unroll
    for (int g = 0; g < kNumGroups; g++) {
      ns[g] = 0;
      npd[g] = 0;
      ch[g] = pr[g] = dp[g] = 0;
      cnt[g] = dsc[g] = qty[g] = 0;
    }
  }
  __device__ __forceinline__ void add(const FastRow& r) {
    const uint32_t s = pack_small(r);
    const uint64_t pd = (uint64_t)disc_prod(r) << 32 | r.price;
    const uint64_t c = fast_charge(r);
    const uint64_t pass_mask = ballot(r.pass);
#pragma unroll
    for (int g = 0; g < kNumGroups; g++) {
      const bool mine = r.pass && r.gid == (uint32_t)g;
      if (ballot(r.gid == (uint32_t)g) & pass_mask) {
        // The empty asm keeps this a separate, wavefront-uniform branch (it cannot be
        // folded into the lane mask), so a group no lane holds costs no adds at all.
        __asm__ volatile("");
        if (mine) {
          ns[g] += s;
          npd[g] += pd;
          ch[g] += c;
        }
      }
    }
  }
  // Called after at most 16 rows: moves the narrow sums into the wide ones.
  __device__ __forceinline__ void flush(bool = false) {
#pragma unroll
    for (int g = 0; g < kNumGroups; g++) {
      if (ballot(ns[g] != 0)) {
        cnt[g] += ns[g] & 31u;
        dsc[g] += (ns[g] >> 5) & 255u;
        qty[g] += ns[g] >> 13;
        pr[g] += (uint32_t)npd[g];
        dp[g] += npd[g] >> 32;
        ns[g] = 0;
        npd[g] = 0;
      }
    }
  }
  __device__ __forceinline__ void totals(LaneTotals& t, char*) const {
#pragma unroll
    for (int g = 0; g < kNumGroups; g++) {
      t.v[g][F_QTY] = qty[g];
      t.v[g][F_PRICE] = pr[g];
      t.v[g][F_DISC] = dsc[g];
      t.v[g][F_CNT] = cnt[g];
      t.v[g][F_DPROD] = dp[g];
      t.v[g][F_CHARGE] = ch[g];
    }
  }
  static constexpr uint32_t lds_bytes(uint32_t) { return 0; }
};

// Packed (count | discount | quantity) in 64 bits for the per-lane 64-bit tables:
// count bits [0,17), discount [17,37), quantity [37,64).  Exact while a lane's table
// sees at most 16384 rows (tuples_per_thread is capped there).
__device__ __forceinline__ uint64_t pack_small64(const FastRow& r) {
  return 1ull | (uint64_t)r.disc << 17 | (uint64_t)r.qty << 37;
}
__device__ __forceinline__ void unpack_small64(uint64_t s, uint64_t* out) {
  out[F_CNT] = s & ((1ull << 17) - 1);
  out[F_DISC] = (s >> 17) & ((1ull << 20) - 1);
  out[F_QTY] = s >> 37;
}
// The narrow per-tile word (pack_small: count [0,5), discount [5,13), quantity [13,32),
// at most 16 rows) re-packed in the pack_small64 layout.
__device__ __forceinline__ uint64_t small_to_small64(uint32_t s) {
  return (uint64_t)(s & 31u) | (uint64_t)((s >> 5) & 255u) << 17 | (uint64_t)(s >> 13) << 37;
}

// Q1_PLACE_IN_REGISTERS_PER_THREAD: the per-thread table of the reference
// (src/kernels/ht_in_registers_per_thread.cuh:54-64): full-width 64-bit sums per group in
// the lane's registers, updated by every row.  The reference selects the row's group with
// a compare-select over all six groups and every field; here each group is one
// wavefront-uniform branch taken only when some lane of the wavefront holds a passing row
// of it (real data fills four of the six groups), and its four 64-bit adds run under the
// lanes' mask: count, discount and quantity share one word (pack_small64).
struct AccPerThread {
  static constexpr bool kPackedSmall = false;
  uint64_t sm[kNumGroups], pr[kNumGroups], dp[kNumGroups], ch[kNumGroups];
  __device__ __forceinline__ void init(char*) {
#pragma unroll
    for (int g = 0; g < kNumGroups; g++) sm[g] = pr[g] = dp[g] = ch[g] = 0;
  }
  __device__ __forceinline__ void add(const FastRow& r) {
    const uint64_t s = pack_small64(r);
    const uint64_t dprod = disc_prod(r);
    const uint64_t c = fast_charge(r);
    const uint64_t pass_mask = ballot(r.pass);
#pragma unroll
    for (int g = 0; g < kNumGroups; g++) {
      if (ballot(r.gid == (uint32_t)g) & pass_mask) {
        // a real, wavefront-uniform branch (not folded into the lane mask), as in AccRegisters
        __asm__ volatile("");
        if (r.pass && r.gid == (uint32_t)g) {
          sm[g] += s;
          pr[g] += r.price;
          dp[g] += dprod;
          ch[g] += c;
        }
      }
    }
  }
  __device__ __forceinline__ void flush(bool = false) {}
  __device__ __forceinline__ void totals(LaneTotals& t, char*) const {
#pragma unroll
    for (int g = 0; g < kNumGroups; g++) {
      unpack_small64(sm[g], t.v[g]);
      t.v[g][F_PRICE] = pr[g];
      t.v[g][F_DPROD] = dp[g];
      t.v[g][F_CHARGE] = ch[g];
    }
  }
  static constexpr uint32_t lds_bytes(uint32_t) { return 0; }
};

// Q1_PLACE_LOCAL_MEM: the per-thread table of the reference
// (src/kernels/ht_in_local_mem.cuh:54-59) lives in the lane's private (scratch) memory:
// [group][count|discount|quantity, price, discount product, charge], 64-bit words.  Rows
// are first summed over kLocalFlushRows rows of the lane in registers, with the
// wavefront-uniform group skip of AccRegisters; those sums then go into the private table:
// the first time as plain stores (the table is never zero-filled), later as one
// read-modify-write of the whole table, and once more after the lane's last row unless
// the last periodic update just happened.  A row costs registers only.  The private tables
// of the resident wavefronts (192 B per lane) outgrow the L2, so their updates reach HBM;
// the update interval sets their cost at 600M compressed rows: every 16 rows 3.2 ms,
// 64 rows 1.46 ms, 256 rows 1.35 ms, 1024 rows 1.19 ms.
constexpr uint32_t kLocalFlushRows = 1024;
struct AccLocal {
  static constexpr bool kPackedSmall = false;
  // Per group: narrow sums of at most 16 rows as in AccRegisters (pack_small word; price
  // and discount product in one 64-bit word), folded every 16 rows into 64-bit
  // register sums of at most kLocalFlushRows rows (count, discount and quantity in the
  // pack_small64 layout, exact up to 16384 rows; price; discount product), with the charge
  // summed in 64 bits directly.
  static constexpr uint32_t kTableWords = kNumGroups * 4;
  static constexpr uint32_t kFlushCalls = kLocalFlushRows / 16;
  uint32_t ns[kNumGroups];
  uint64_t npd[kNumGroups];
  uint64_t sm[kNumGroups], pr[kNumGroups], dp[kNumGroups], ch[kNumGroups];
  uint32_t calls;
  // The table (kTableWords words of the caller's private array) is used only through this
  // pointer, which an empty asm hides from the optimizer: the array cannot become
  // registers, so it stays in private memory, and its loads and stores are ordinary ones
  // the compiler may overlap (volatile accesses would each wait for the previous one).
  uint64_t* t;
  __device__ __forceinline__ void init(char* private_table) {
    calls = 0;
    t = reinterpret_cast<uint64_t*>(private_table);
    __asm__ volatile("" : "+v"(t));
#pragma unroll
    for (int g = 0; g < kNumGroups; g++) {
      ns[g] = 0;
      npd[g] = 0;
      sm[g] = pr[g] = dp[g] = ch[g] = 0;
    }
  }
  __device__ __forceinline__ void add(const FastRow& r) {
    const uint32_t s = pack_small(r);
    const uint64_t pd = (uint64_t)disc_prod(r) << 32 | r.price;
    const uint64_t c = fast_charge(r);
    const uint64_t pass_mask = ballot(r.pass);
#pragma unroll
    for (int g = 0; g < kNumGroups; g++) {
      if (ballot(r.gid == (uint32_t)g) & pass_mask) {
        __asm__ volatile("");
        if (r.pass && r.gid == (uint32_t)g) {
          ns[g] += s;
          npd[g] += pd;
          ch[g] += c;
        }
      }
    }
  }
  // Called every 16 rows of a lane (at the same points in every lane; a lane's rows are a
  // whole number of 16-row periods, so the last such call comes right before the final
  // one) and after its last row (last = true).  Every call folds the narrow sums into the
  // register sums; every kLocalFlushRows rows, and at the end unless that just happened,
  // the register sums go into the private table.
  __device__ __forceinline__ void flush(bool last = false) {
    static_assert(kLocalFlushRows <= 16384, "pack_small64 is exact up to 16384 rows");
    if (last) {
      if (calls > 0 && calls % kFlushCalls == 0) return;  // nothing summed since the last update
    } else {
#pragma unroll
      for (int g = 0; g < kNumGroups; g++) {
        if (ballot(ns[g] != 0)) {
          sm[g] += small_to_small64(ns[g]);
          pr[g] += (uint32_t)npd[g];
          dp[g] += npd[g] >> 32;
          ns[g] = 0;
          npd[g] = 0;
        }
      }
      if (++calls % kFlushCalls != 0) return;
    }
    // The table is updated in two halves of three groups: each half's words are loaded
    // before any is added to and stored, with no branch between the loads (groups without
    // rows add zero), so the lane waits for its private memory twice per update while
    // holding only half the table in registers.
    const bool first = calls <= kFlushCalls;
    constexpr int kHalf = kNumGroups / 2;
#pragma unroll
    for (int h = 0; h < kNumGroups; h += kHalf) {
      uint64_t w[kHalf * 4];
#pragma unroll
      for (int i = 0; i < kHalf * 4; i++) w[i] = first ? 0 : t[h * 4 + i];
#pragma unroll
      for (int g = 0; g < kHalf; g++) {
        w[g * 4 + 0] += sm[h + g];
        w[g * 4 + 1] += pr[h + g];
        w[g * 4 + 2] += dp[h + g];
        w[g * 4 + 3] += ch[h + g];
        sm[h + g] = pr[h + g] = dp[h + g] = ch[h + g] = 0;
      }
#pragma unroll
      for (int i = 0; i < kHalf * 4; i++) t[h * 4 + i] = w[i];
    }
  }
  // The private table (after the last flush).
  __device__ __forceinline__ void totals(LaneTotals& o, char*) const {
#pragma unroll
    for (int g = 0; g < kNumGroups; g++) {
      const uint64_t* w = t + g * 4;
      unpack_small64(w[0], o.v[g]);
      o.v[g][F_PRICE] = w[1];
      o.v[g][F_DPROD] = w[2];
      o.v[g][F_CHARGE] = w[3];
    }
  }
  static constexpr uint32_t lds_bytes(uint32_t) { return 0; }
};

// Q1_PLACE_PER_THREAD_SHARED_MEM: each lane owns a table in LDS laid out
// [group][field][thread] (consecutive lanes -> consecutive 8-byte words, so a wavefront's
// update of any mix of groups is bank-conflict free); rows are added with return-less
// LDS atomics, which need no VALU add and no read-back.  With the block size known at
// compile time a row's address is one shift-add and the four fields are immediate offsets.
constexpr uint32_t shared_table_bytes(uint32_t threads) { return 4u * kNumGroups * 8u * threads; }

// NARROW: at most 256 rows per lane, so a lane's price sum stays below 256 * 2^24 = 2^32
// and the price field takes a 32-bit LDS atomic on the low word of its 64-bit slot (no
// register pair to build per row); the slot's high word stays zero.
template <int B, bool NARROW = false>
struct AccShared {
  uint64_t* tab;  // dynamic LDS: 6 * 4 * B words
  __device__ __forceinline__ void init(char* lds) {
    tab = reinterpret_cast<uint64_t*>(lds);
#pragma unroll
    for (int i = 0; i < 4 * kNumGroups; i++) tab[i * B + thread_index()] = 0;
  }
  __device__ __forceinline__ void add(const FastRow& r) {
    // Nearly every row passes: lay the atomics out inline, with no branch around them.
    if (__builtin_expect(r.counts(), 1)) {
      uint64_t* base = tab + r.gid * (4 * B) + thread_index();
      __hip_atomic_fetch_add(base, pack_small64(r), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (NARROW)
        __hip_atomic_fetch_add(reinterpret_cast<uint32_t*>(base + B), r.price, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
      else
        __hip_atomic_fetch_add(base + B, (uint64_t)r.price, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_fetch_add(base + 2 * B, (uint64_t)disc_prod(r), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_fetch_add(base + 3 * B, fast_charge(r), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
  __device__ __forceinline__ void flush(bool = false) {}
  // NARROW tables hand over the packed (count | discount | quantity) word in F_CNT: at
  // most 256 rows per lane, a wavefront's sum of it still fits every field (count
  // < 2^17, discount < 2^20, quantity < 2^27), so the block merge sums one word for three.
  static constexpr bool kPackedSmall = NARROW;
  __device__ __forceinline__ void totals(LaneTotals& o, char*) const {
#pragma unroll
    for (int g = 0; g < kNumGroups; g++) {
      const uint64_t* base = tab + g * (4 * B) + thread_index();
      if (NARROW) {
        o.v[g][F_CNT] = base[0];
        o.v[g][F_QTY] = o.v[g][F_DISC] = 0;
      } else {
        unpack_small64(base[0], o.v[g]);
      }
      o.v[g][F_PRICE] = base[B];
      o.v[g][F_DPROD] = base[2 * B];
      o.v[g][F_CHARGE] = base[3 * B];
    }
  }
  static constexpr uint32_t lds_bytes(uint32_t) { return shared_table_bytes(B); }
};

// Variant of the LDS table for compressed-layout launches of at most 256 rows per lane
// (quantities in whole units, at most 63): three slots per group instead of four.
//   slot 0: price sum (< 256 * 2^24 = 2^32) in bits [0,32), discount sum (<= 256 * 15
//           < 2^12) in [32,44), quantity sum (<= 256 * 63 < 2^14) in [44,58);
//   slot 1: discount-price product sum (< 256 * 15 * 2^24 < 2^36) in [0,36), count
//           (<= 256) in [48,57);
//   slot 2: charge.
// Three return-less 64-bit LDS atomics per passing row instead of four, and a table of
// 6 * 3 * 8 bytes per lane, so four 256-thread blocks fit a compute unit's LDS.
constexpr uint32_t packed_table_bytes(uint32_t threads) { return 3u * kNumGroups * 8u * threads; }

template <int B>
struct AccSharedPacked {
  uint64_t* tab;  // dynamic LDS: 6 * 3 * B words
  __device__ __forceinline__ void init(char* lds) {
    tab = reinterpret_cast<uint64_t*>(lds);
#pragma unroll
    for (int i = 0; i < 3 * kNumGroups; i++) tab[i * B + thread_index()] = 0;
  }
  __device__ __forceinline__ void add(const FastRow& r) {
    // Nearly every row passes: lay the atomics out inline, with no branch around them.
    if (__builtin_expect(r.counts(), 1)) {
      uint64_t* base = tab + r.gid * (3 * B) + thread_index();
      const uint64_t w0 = (uint64_t)r.price | (uint64_t)(r.disc | r.qty << 12) << 32;
      const uint64_t w1 = (uint64_t)disc_prod(r) | 1ull << 48;
      __hip_atomic_fetch_add(base, w0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_fetch_add(base + B, w1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_fetch_add(base + 2 * B, fast_charge(r), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
  __device__ __forceinline__ void flush(bool = false) {}
  // The block merge gets count, discount and quantity re-packed in the pack_small64 layout
  // (a wavefront's sums of them fit it: count <= 64 * 256, discount <= 64 * 3840,
  // quantity <= 64 * 16128), so it sums one word for the three.
  static constexpr bool kPackedSmall = true;
  __device__ __forceinline__ void totals(LaneTotals& o, char*) const {
#pragma unroll
    for (int g = 0; g < kNumGroups; g++) {
      const uint64_t* base = tab + g * (3 * B) + thread_index();
      const uint64_t w0 = base[0], w1 = base[B];
      o.v[g][F_PRICE] = w0 & 0xFFFFFFFFull;
      o.v[g][F_CNT] = (w1 >> 48) | ((w0 >> 32) & 0xFFFull) << 17 | (w0 >> 44) << 37;
      o.v[g][F_QTY] = o.v[g][F_DISC] = 0;
      o.v[g][F_DPROD] = w1 & ((1ull << 48) - 1);
      o.v[g][F_CHARGE] = base[2 * B];
    }
  }
  static constexpr uint32_t lds_bytes(uint32_t) { return packed_table_bytes(B); }
};

// Q1_PLACE_GLOBAL: the block's aggregation table lives in global memory (the kernel's
// workspace, one table per block), as the reference's table does
// (src/kernels/ht_in_global_mem.hpp:12-52).  The reference adds every row into it with
// five device atomics; here rows are summed in the lanes' registers (AccRegisters), each
// wavefront reduces its lanes' sums with DPP and adds them into the block's table with
// one atomic per (group, field) from one lane, and the block then merges its table into
// the destination once: thread g reads group g's words and issues its destination atomics.
// The table sees up to blockDim * 16384 <= 2^22 rows: count and discount share one word
// (count bits [0,24), discount [24,64)), the other sums have a word each; every block sum
// fits 64 bits (price < 2^46, discount product < 2^50, charge < 2^60).
__device__ __forceinline__ void atomic_add128(uint64_t* lo_p, uint64_t* hi_p, uint64_t lo, uint64_t hi);
struct AccGlobal {
  static constexpr bool kPackedSmall = false;
  static constexpr int W = 5;  // words per group: count|disc, qty, price, dprod, charge
  AccRegisters regs;
  uint64_t* tab;               // [W][6] words of this block
  __device__ __forceinline__ void init_global(uint64_t* ws) {
    tab = ws + (uint64_t)blockIdx.x * (W * kNumGroups);
    if (thread_index() < W * kNumGroups)
      __hip_atomic_store(tab + thread_index(), 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
  }
  __device__ __forceinline__ void init(char* lds) { regs.init(lds); }
  __device__ __forceinline__ void add(const FastRow& r) { regs.add(r); }
  __device__ __forceinline__ void flush(bool = false) { regs.flush(); }
  // The block's whole merge (called by every thread of the block).
  __device__ __forceinline__ void merge(q1_device_aggregates* out, uint64_t qty_scale) const {
    LaneTotals o;
    regs.totals(o, nullptr);
    const uint32_t lane = thread_index() & 63;
#pragma unroll
    for (int g = 0; g < kNumGroups; g++) {
      if (ballot(o.v[g][F_CNT] != 0)) {
        // a lane's count is at most 16384 and its discount sum at most 15 * 16384
        const uint64_t cd = wave_sum_u64(o.v[g][F_CNT] | o.v[g][F_DISC] << 24);
        const uint64_t q = wave_sum_u64(o.v[g][F_QTY]);
        const uint64_t p = wave_sum_u64(o.v[g][F_PRICE]);
        const uint64_t d = wave_sum_u64(o.v[g][F_DPROD]);
        const uint64_t c = wave_sum_u64(o.v[g][F_CHARGE]);
        if (lane == 0) {
          atomicAdd((unsigned long long*)(tab + g), (unsigned long long)cd);
          atomicAdd((unsigned long long*)(tab + kNumGroups + g), (unsigned long long)q);
          atomicAdd((unsigned long long*)(tab + 2 * kNumGroups + g), (unsigned long long)p);
          atomicAdd((unsigned long long*)(tab + 3 * kNumGroups + g), (unsigned long long)d);
          atomicAdd((unsigned long long*)(tab + 4 * kNumGroups + g), (unsigned long long)c);
        }
      }
    }
    // Only this block reads its table back, and every wave of the block shares one L2 (where
    // the atomics are performed): the barrier's workgroup-scope release and acquire order
    // them before the read-back.  A device-scope fence here would make each block wait for
    // an L2 write-back towards the other XCDs.
    __syncthreads();
    if (thread_index() < kNumGroups) {
      const int g = thread_index();
      auto ld = [&](int w) {
        return __hip_atomic_load(tab + w * kNumGroups + g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      };
      const uint64_t cd = ld(0), q = ld(1), p = ld(2), d = ld(3), c = ld(4);
      if (cd) {
        atomicAdd((unsigned long long*)&out->sum_qty[g], (unsigned long long)(q * qty_scale));
        atomicAdd((unsigned long long*)&out->sum_base_price[g], (unsigned long long)p);
        atomicAdd((unsigned long long*)&out->sum_disc[g], (unsigned long long)(cd >> 24));
        atomicAdd((unsigned long long*)&out->count[g], (unsigned long long)(cd & ((1ull << 24) - 1)));
        // sum_disc_price = 100 * sum_price - sum(disc * price) >= 0 (every disc <= 15)
        atomic_add128(&out->sum_disc_price_lo[g], &out->sum_disc_price_hi[g], 100 * p - d, 0);
        atomic_add128(&out->sum_charge_lo[g], &out->sum_charge_hi[g], c, 0);
      }
    }
  }
  static constexpr uint32_t lds_bytes(uint32_t) { return 0; }
};

template <int P, int B> struct AccOf;
template <int B> struct AccOf<Q1_PLACE_IN_REGISTERS, B> { using A = AccRegisters; };
template <int B> struct AccOf<Q1_PLACE_IN_REGISTERS_PER_THREAD, B> { using A = AccPerThread; };
template <int B> struct AccOf<Q1_PLACE_LOCAL_MEM, B> { using A = AccLocal; };
template <int B> struct AccOf<Q1_PLACE_PER_THREAD_SHARED_MEM, B> { using A = AccShared<B>; };
template <int B> struct AccOf<Q1_PLACE_GLOBAL, B> { using A = AccGlobal; };
// Internal kernel variant of the shared placement for launches of at most 256 rows per lane.
constexpr int kPlaceSharedNarrow = 8;
template <int B> struct AccOf<kPlaceSharedNarrow, B> { using A = AccShared<B, true>; };
// ... and, for the compressed layouts, the packed table with three slots per group.
constexpr int kPlaceSharedPacked = 9;
template <int B> struct AccOf<kPlaceSharedPacked, B> { using A = AccSharedPacked<B>; };

template <> struct ScanTileOf<Q1_LAYOUT_COMPRESSED, Q1_PLACE_PER_THREAD_SHARED_MEM> { using T = CompressedStreamTile<false>; };
template <> struct ScanTileOf<Q1_LAYOUT_COMPRESSED, kPlaceSharedNarrow> { using T = CompressedStreamTile<false>; };
template <> struct ScanTileOf<Q1_LAYOUT_COMPRESSED_PUSHDOWN, Q1_PLACE_PER_THREAD_SHARED_MEM> {
  using T = CompressedStreamTile<true>;
};
template <> struct ScanTileOf<Q1_LAYOUT_COMPRESSED_PUSHDOWN, kPlaceSharedNarrow> { using T = CompressedStreamTile<true>; };
template <> struct ScanTileOf<Q1_LAYOUT_UNCOMPRESSED, Q1_PLACE_PER_THREAD_SHARED_MEM> { using T = UncompressedStreamTile; };
template <> struct ScanTileOf<Q1_LAYOUT_UNCOMPRESSED, kPlaceSharedNarrow> { using T = UncompressedStreamTile; };
template <> struct ScanTileOf<Q1_LAYOUT_REFERENCE, Q1_PLACE_PER_THREAD_SHARED_MEM> { using T = ReferenceStreamTile; };
template <> struct ScanTileOf<Q1_LAYOUT_REFERENCE, kPlaceSharedNarrow> { using T = ReferenceStreamTile; };
template <int L> struct ScanTileOf<L, kPlaceSharedPacked> { using T = typename ScanTileOf<L, kPlaceSharedNarrow>::T; };

// ------------------------------------------------------------------------------------
// Block merge and the single global merge per block.
// ------------------------------------------------------------------------------------
// 128-bit signed add of (lo, hi) into a device (lo, hi) word pair with two atomics.
__device__ __forceinline__ void atomic_add128(uint64_t* lo_p, uint64_t* hi_p, uint64_t lo, uint64_t hi) {
  uint64_t carry = 0;
  if (lo) {
    const uint64_t old = atomicAdd((unsigned long long*)lo_p, (unsigned long long)lo);
    carry = (old + lo) < old;
  }
  if (hi + carry) atomicAdd((unsigned long long*)hi_p, (unsigned long long)(hi + carry));
}

// Block-wide reduction of LaneTotals (exact: lane sums fit uint64, block sums are taken
// in 128 bits) and the merge into `out`.  `red` is LDS scratch of
// (blockDim/64) * 6 * NF uint64 words.  qty_scale turns layout units into x100.
// PACKED: t.v[g][F_CNT] holds the packed (count | discount | quantity) word of the
// lane (pack_small64 layout) and F_QTY / F_DISC are unused.
template <bool PACKED = false>
__device__ void block_merge(const LaneTotals& t, uint64_t* red, q1_device_aggregates* out,
                            uint64_t qty_scale) {
  const uint32_t wave = thread_index() >> 6, lane = thread_index() & 63, nwaves = blockDim.x >> 6;
#pragma unroll
  for (int g = 0; g < kNumGroups; g++) {
    const bool any = ballot(t.v[g][F_CNT] != 0) != 0;
    uint64_t* r = red + (wave * kNumGroups + g) * NF;
#pragma unroll
    for (int f = 0; f < NF; f++) {
      if (PACKED && (f == F_QTY || f == F_DISC)) continue;
      const uint64_t s = any ? wave_sum_u64(t.v[g][f]) : 0;
      if (lane == 0) {
        if (PACKED && f == F_CNT) {
          uint64_t u[NF];
          unpack_small64(s, u);
          r[F_CNT] = u[F_CNT];
          r[F_DISC] = u[F_DISC];
          r[F_QTY] = u[F_QTY];
        } else {
          r[f] = s;
        }
      }
    }
  }
  __syncthreads();
  // Thread g*NF + f sums field f of group g over the waves (128-bit) and issues that
  // field's atomics, so a block's destination atomics go out together.
  if (thread_index() < kNumGroups * NF) {
    const int g = thread_index() / NF, f = thread_index() % NF;
    uint64_t lo = 0, hi = 0, cnt = 0, plo = 0, phi = 0;
    for (uint32_t w = 0; w < nwaves; w++) {
      const uint64_t* r = red + (w * kNumGroups + g) * NF;
      const uint64_t v = r[f];
      lo += v;
      hi += lo < v;
      cnt |= r[F_CNT];
      const uint64_t pv = r[F_PRICE];
      plo += pv;
      phi += plo < pv;
    }
    if (cnt) {
      switch (f) {
        case F_QTY: atomicAdd((unsigned long long*)&out->sum_qty[g], (unsigned long long)(lo * qty_scale)); break;
        case F_PRICE: atomicAdd((unsigned long long*)&out->sum_base_price[g], (unsigned long long)lo); break;
        case F_DISC: atomicAdd((unsigned long long*)&out->sum_disc[g], (unsigned long long)lo); break;
        case F_CNT: atomicAdd((unsigned long long*)&out->count[g], (unsigned long long)lo); break;
        case F_DPROD: {
          // sum_disc_price = 100 * sum_price - sum(disc * price), in 128 bits
          const unsigned __int128 p100 = ((unsigned __int128)phi << 64 | plo) * 100u;
          const unsigned __int128 dp = p100 - ((unsigned __int128)hi << 64 | lo);
          atomic_add128(&out->sum_disc_price_lo[g], &out->sum_disc_price_hi[g], (uint64_t)dp, (uint64_t)(dp >> 64));
          break;
        }
        default: atomic_add128(&out->sum_charge_lo[g], &out->sum_charge_hi[g], lo, hi); break;
      }
    }
  }
}

// ------------------------------------------------------------------------------------
// Row bodies of one tile group.
// ------------------------------------------------------------------------------------
// A whole group inside the range.  Rows with the unused return-flag code are first made
// to fail the filter (a wavefront-uniform branch that real data never takes), so every
// row's slot is below 6 and the slot test drops out of the row body.
template <class Tile, int U, class Acc>
__device__ __forceinline__ void add_whole_group(TileGroup<Tile, U>& grp, Acc& acc, int32_t cut) {
  constexpr int TILE = Tile::TILE;
  uint32_t bad = 0;
#pragma unroll
  for (int m = 0; m < U; m++) bad |= grp.tile[m].bad_codes();
  if (ballot(bad != 0)) {
#pragma unroll
    for (int m = 0; m < U; m++) grp.tile[m].drop_bad_codes();
  }
#pragma unroll
  for (int m = 0; m < U; m++)
#pragma unroll
    for (int j = 0; j < TILE; j++) {
      const FastRow r = grp.tile[m].template fast<false>(j, cut, 0);
      __builtin_assume(r.gid < kNumGroups);
      acc.add(r);
    }
}

// Any group: tiles may be unused, or partial at the range's ends.
template <class Tile, int U, class Acc>
__device__ __forceinline__ void add_any_group(const TileGroup<Tile, U>& grp, Acc& acc, int32_t cut) {
  constexpr int TILE = Tile::TILE;
  if (grp.plain) {
#pragma unroll
    for (int m = 0; m < U; m++)
#pragma unroll
      for (int j = 0; j < TILE; j++) acc.add(grp.tile[m].template fast<false>(j, cut, 0));
  } else {
#pragma unroll
    for (int m = 0; m < U; m++) {
      if (!grp.use[m]) continue;
      if (!grp.edge[m]) {
#pragma unroll
        for (int j = 0; j < TILE; j++) acc.add(grp.tile[m].template fast<false>(j, cut, 0));
      } else {
#pragma unroll
        for (int j = 0; j < TILE; j++) acc.add(grp.tile[m].template fast<true>(j, cut, grp.valid[m]));
      }
    }
  }
}

// ------------------------------------------------------------------------------------
// The fast scan kernel: layout L, placement P.
// ------------------------------------------------------------------------------------
// One block's share of a range: block `block` of the grid that launch_scan would run over
// a's rows, merged into *out.  `lds` is the block's dynamic LDS.
template <int L, int P, int B>
__device__ __forceinline__ void scan_block(const ScanArgs& a, uint64_t block, uint32_t host_tiles_per_thread,
                                           uint64_t* ws, uint64_t* lds, q1_device_aggregates* out) {
  using Tile = typename ScanTileOf<L, P>::T;
  using Acc = typename AccOf<P, B>::A;
  constexpr int TILE = Tile::TILE;
  constexpr int U = Tile::GROUP;                       // tiles loaded together per lane
  constexpr int FLUSH = U * TILE >= 16 ? 1 : 16 / (U * TILE);  // groups per 16 rows
  // The host counts tiles of TileOf<L> (tile_rows); a smaller kernel tile splits each.
  constexpr int SUB = TileOf<L>::T::TILE / TILE;

  const int32_t cut = column_cutoff<L>(a.cutoff_day);
  // Block b covers kernel tiles [first_tile + b*B*tpt, +B*tpt): SUB times the host's
  // tiles per lane, so the grid covers every row of the range.  The first tile is that of
  // row_begin rounded down to kRowAlign rows, as the host counts them.
  static_assert(kRowAlign % TILE == 0, "kRowAlign is a whole number of tiles");
  const uint32_t tiles_per_thread = host_tiles_per_thread * SUB;
  const uint64_t first_tile = a.row_begin / kRowAlign * kRowAlign / TILE;
  const uint64_t end_tile = (a.row_end + TILE - 1) / TILE;
  const uint64_t base = first_tile + block * B * tiles_per_thread;

  Acc acc;
  // the local placement's table in the lane's private memory (unused by the others)
  uint64_t private_table[P == Q1_PLACE_LOCAL_MEM ? AccLocal::kTableWords : 1];
  if constexpr (P == Q1_PLACE_GLOBAL) acc.init_global(ws);
  if constexpr (P == Q1_PLACE_LOCAL_MEM)
    acc.init(reinterpret_cast<char*>(private_table));
  else
    acc.init(reinterpret_cast<char*>(lds));

  constexpr bool kSplit = P == Q1_PLACE_PER_THREAD_SHARED_MEM || P == kPlaceSharedNarrow || P == kPlaceSharedPacked;
  if constexpr (kSplit) {
    // Groups k in [k_lo, k_end) are whole and inside the range for every lane of the
    // block: the tiles of group k are base + (k + m) * B + lane, and they all lie in
    // [whole_begin, whole_end), the tiles with no row outside [row_begin, row_end).  They
    // run a branch-free body; the few others (the range's first and last groups) a
    // guarded one.  A dropped row's ship date 0xFFFF must fail the filter: compressed
    // cutoffs of 65535 and above (beyond the year 2170) take the guarded body throughout.
    const uint64_t whole_begin = (a.row_begin + TILE - 1) / TILE, whole_end = a.row_end / TILE;
    uint32_t k_lo = 0, k_end = 0;
    if ((L != Q1_LAYOUT_COMPRESSED || cut < 65535) && whole_end > base) {
      const uint64_t lo = base >= whole_begin ? 0 : (whole_begin - base + B - 1) / B;
      const uint64_t klo = (lo + U - 1) / U * U;
      uint64_t kmax = (whole_end - base) / B;  // groups with k + U <= kmax are inside
      if (kmax > tiles_per_thread) kmax = tiles_per_thread;
      if (kmax >= klo + U) {
        k_lo = (uint32_t)klo;
        k_end = (uint32_t)(klo + (kmax - klo) / U * U);
      }
    }
    uint32_t k = 0, n = 0;
    for (; k < k_lo; k += U, n++) {
      TileGroup<Tile, U> grp;
      grp.load(a.cols, base + (uint64_t)k * B + thread_index(), B, tiles_per_thread - k, end_tile, a.row_begin,
               a.row_end);
      add_any_group(grp, acc, cut);
      if ((n + 1) % FLUSH == 0) acc.flush();  // at most 16 rows per lane between flushes
    }
    for (; k < k_end; k += U, n++) {
      TileGroup<Tile, U> grp;
      grp.load_plain(a.cols, base + (uint64_t)k * B + thread_index(), B);
      add_whole_group(grp, acc, cut);
      if ((n + 1) % FLUSH == 0) acc.flush();
    }
    for (; k < tiles_per_thread; k += U, n++) {
      TileGroup<Tile, U> grp;
      grp.load(a.cols, base + (uint64_t)k * B + thread_index(), B, tiles_per_thread - k, end_tile, a.row_begin,
               a.row_end);
      add_any_group(grp, acc, cut);
      if ((n + 1) % FLUSH == 0) acc.flush();
    }
  } else {
    // The register-table placements keep one guarded loop: a second copy of the row body
    // costs them registers (spills, lower occupancy) that the saved instructions do not
    // repay.
    for (uint32_t k = 0, n = 0; k < tiles_per_thread; k += U, n++) {
      TileGroup<Tile, U> grp;
      grp.load(a.cols, base + (uint64_t)k * B + thread_index(), B, tiles_per_thread - k, end_tile, a.row_begin,
               a.row_end);
      if (grp.plain) {
#pragma unroll
        for (int m = 0; m < U; m++)
#pragma unroll
          for (int j = 0; j < TILE; j++) acc.add(grp.tile[m].template fast<false>(j, cut, 0));
      } else {
#pragma unroll
        for (int m = 0; m < U; m++) {
          if (!grp.use[m]) continue;
          if (!grp.edge[m]) {
#pragma unroll
            for (int j = 0; j < TILE; j++) acc.add(grp.tile[m].template fast<false>(j, cut, 0));
          } else {
#pragma unroll
            for (int j = 0; j < TILE; j++) acc.add(grp.tile[m].template fast<true>(j, cut, grp.valid[m]));
          }
        }
      }
      if ((n + 1) % FLUSH == 0) acc.flush();  // at most 16 rows per lane between flushes
    }
  }
  acc.flush(true);

  const uint64_t qty_scale =
      (L == Q1_LAYOUT_COMPRESSED || L == Q1_LAYOUT_COMPRESSED_PUSHDOWN) ? kQuantityScaleCompressed : 1;
  if constexpr (P == Q1_PLACE_GLOBAL) {
    // through the block's table in global memory
    acc.merge(out, qty_scale);
  } else {
    LaneTotals tot;
    acc.totals(tot, reinterpret_cast<char*>(lds));
    // The reduction scratch follows the placement's own LDS table.
    uint64_t* red = lds + Acc::lds_bytes(B) / 8;
    __syncthreads();
    block_merge<Acc::kPackedSmall>(tot, red, out, qty_scale);
  }
}

template <int L, int P, int B>
__global__ void __launch_bounds__(B)
scan_fast(ScanArgs a, uint32_t host_tiles_per_thread, uint64_t* ws) {
  extern __shared__ uint64_t lds[];
  scan_block<L, P, B>(a, blockIdx.x, host_tiles_per_thread, ws, lds, a.out + (blockIdx.x & a.out_mask));
}

// ------------------------------------------------------------------------------------
// The any-value scan kernel (tables outside the TPC-H domain, and q1_enqueue without
// Q1_LAUNCH_ASSUME_TPCH_DOMAIN): exact for every input with the reference's arithmetic,
// wrapping int64 per-row products and 128-bit signed sums.  Each lane tests its tile:
// a tile whose rows all lie inside the TPC-H value domain goes through the narrow
// per-lane LDS table of the shared placement (exact there: no product can wrap); any
// other tile goes through a per-lane 64-bit register table.  At the end the narrow table
// is folded into the wide one (disc_price = 100 * price - disc * price, all
// non-negative), and the block merges in 128 bits.  In-domain data runs at the speed of
// the narrow kernel; a wavefront pays for the wide path only on tiles that need it.
// ------------------------------------------------------------------------------------
struct Wide128 {
  uint64_t lo;
  uint64_t hi;
  __device__ __forceinline__ void add64s(int64_t v) {
    const uint64_t nlo = lo + (uint64_t)v;
    hi += (uint64_t)(v >> 63) + (nlo < lo);
    lo = nlo;
  }
};

// The sums of one slot for rows outside the TPC-H domain: 64-bit (wrapping, as the
// reference's int64 sums) and 128-bit products.
struct WideSums {
  uint64_t qty, price, disc, cnt;
  Wide128 dp, ch;
};

template <class Tile>
__device__ __forceinline__ void add_wide_rows(const Tile& tile, WideSums* wide, int32_t cut, uint32_t valid) {
#pragma unroll
  for (int j = 0; j < Tile::TILE; j++) {
    const WideRow r = tile.wide(j, cut, valid);
    if (r.gid < kNumGroups) {
      const int64_t disc_1 = (int64_t)(100ull - (uint64_t)r.disc);
      const int64_t disc_price = (int64_t)((uint64_t)disc_1 * (uint64_t)r.price);
      const int64_t charge = (int64_t)((uint64_t)disc_price * (uint64_t)(r.tax + 100));
      WideSums& t = wide[r.gid];
      t.qty += (uint64_t)r.qty;
      t.price += (uint64_t)r.price;
      t.disc += (uint64_t)r.disc;
      t.cnt += 1;
      t.dp.add64s(disc_price);
      t.ch.add64s(charge);
    }
  }
}

// A group of tiles some of whose rows may lie outside the domain: each lane sends each of
// its tiles to the narrow LDS table or to its wide table.
template <class Tile, int U, class Acc>
__device__ __forceinline__ void add_group_any_value(const TileGroup<Tile, U>& grp, Acc& narrow, WideSums* wide,
                                                    bool& wide_used, int32_t cut, bool whole) {
  constexpr int TILE = Tile::TILE;
#pragma unroll
  for (int m = 0; m < U; m++) {
    if (!whole && !grp.use[m]) continue;
    const Tile& tile = grp.tile[m];
    const bool edge = !whole && grp.edge[m];
    const uint32_t valid = whole ? (1u << TILE) - 1 : grp.valid[m];
    if (tile.in_domain()) {
      if (!edge) {
#pragma unroll
        for (int j = 0; j < TILE; j++) narrow.add(tile.template fast<false>(j, cut, 0));
      } else {
#pragma unroll
        for (int j = 0; j < TILE; j++) narrow.add(tile.template fast<true>(j, cut, valid));
      }
    } else {
      if (!wide_used) {  // the lane's first row outside the domain
#pragma unroll
        for (int g = 0; g < kNumGroups; g++) wide[g] = WideSums{0, 0, 0, 0, {0, 0}, {0, 0}};
        wide_used = true;
      }
      add_wide_rows(tile, wide, cut, valid);
    }
  }
}

template <int L, int B>
__global__ void __launch_bounds__(B)
scan_any(ScanArgs a, uint32_t host_tiles_per_thread) {
  // The LDS placement's streaming tiles; the host counts tiles of TileOf<L>.
  using Tile = typename ScanTileOf<L, kPlaceSharedNarrow>::T;
  constexpr int TILE = Tile::TILE;
  constexpr int U = Tile::GROUP;
  constexpr int SUB = TileOf<L>::T::TILE / TILE;
  extern __shared__ uint64_t lds[];

  const uint32_t tiles_per_thread = host_tiles_per_thread * SUB;
  const int32_t cut = column_cutoff<L>(a.cutoff_day);
  const uint64_t first_tile = a.row_begin / kRowAlign * kRowAlign / TILE;  // as the host counts them
  const uint64_t end_tile = (a.row_end + TILE - 1) / TILE;
  const uint64_t base = first_tile + (uint64_t)blockIdx.x * B * tiles_per_thread;

  AccShared<B> narrow;
  narrow.init(reinterpret_cast<char*>(lds));
  // Indexed by slot at run time, so it lives in the lane's private memory and costs no
  // registers; only lanes with rows outside the domain clear, update and read it.
  WideSums wide[kNumGroups];
  bool wide_used = false;

  // Whole groups as in scan_fast: when every row of a wavefront's group lies in the domain
  // it takes the branch-free body; otherwise its lanes sort their tiles.
  const uint64_t whole_begin = (a.row_begin + TILE - 1) / TILE, whole_end = a.row_end / TILE;
  uint32_t k_lo = 0, k_end = 0;
  if ((L != Q1_LAYOUT_COMPRESSED || cut < 65535) && whole_end > base) {
    const uint64_t lo = base >= whole_begin ? 0 : (whole_begin - base + B - 1) / B;
    const uint64_t klo = (lo + U - 1) / U * U;
    uint64_t kmax = (whole_end - base) / B;
    if (kmax > tiles_per_thread) kmax = tiles_per_thread;
    if (kmax >= klo + U) {
      k_lo = (uint32_t)klo;
      k_end = (uint32_t)(klo + (kmax - klo) / U * U);
    }
  }
  uint32_t k = 0;
  for (; k < k_lo; k += U) {
    TileGroup<Tile, U> grp;
    grp.load(a.cols, base + (uint64_t)k * B + thread_index(), B, tiles_per_thread - k, end_tile, a.row_begin,
             a.row_end);
    add_group_any_value(grp, narrow, wide, wide_used, cut, false);
  }
  for (; k < k_end; k += U) {
    TileGroup<Tile, U> grp;
    grp.load_plain(a.cols, base + (uint64_t)k * B + thread_index(), B);
    bool in = true;
#pragma unroll
    for (int m = 0; m < U; m++) in = in & grp.tile[m].in_domain();
    if (ballot(!in) == 0)
      add_whole_group(grp, narrow, cut);
    else
      add_group_any_value(grp, narrow, wide, wide_used, cut, true);
  }
  for (; k < tiles_per_thread; k += U) {
    TileGroup<Tile, U> grp;
    grp.load(a.cols, base + (uint64_t)k * B + thread_index(), B, tiles_per_thread - k, end_tile, a.row_begin,
             a.row_end);
    add_group_any_value(grp, narrow, wide, wide_used, cut, false);
  }

  uint64_t qty[kNumGroups], price[kNumGroups], disc[kNumGroups], cnt[kNumGroups];
  Wide128 dp[kNumGroups], ch[kNumGroups];
#pragma unroll
  for (int g = 0; g < kNumGroups; g++) {
    qty[g] = price[g] = disc[g] = cnt[g] = 0;
    dp[g].lo = dp[g].hi = ch[g].lo = ch[g].hi = 0;
  }
  if (wide_used) {
#pragma unroll
    for (int g = 0; g < kNumGroups; g++) {
      qty[g] = wide[g].qty;
      price[g] = wide[g].price;
      disc[g] = wide[g].disc;
      cnt[g] = wide[g].cnt;
      dp[g] = wide[g].dp;
      ch[g] = wide[g].ch;
    }
  }

  // Fold the lane's narrow table into its wide one.  Lane sums of the narrow table are
  // exact and far below 2^63 (at most 16384 rows per lane).
  {
    LaneTotals nt;
    narrow.totals(nt, reinterpret_cast<char*>(lds));
#pragma unroll
    for (int g = 0; g < kNumGroups; g++) {
      qty[g] += nt.v[g][F_QTY];
      price[g] += nt.v[g][F_PRICE];
      disc[g] += nt.v[g][F_DISC];
      cnt[g] += nt.v[g][F_CNT];
      dp[g].add64s((int64_t)(100u * nt.v[g][F_PRICE] - nt.v[g][F_DPROD]));
      ch[g].add64s((int64_t)nt.v[g][F_CHARGE]);
    }
  }

  // Block merge in 128 bits for every field; the scratch follows the narrow table.
  const uint32_t wave = thread_index() >> 6, lane = thread_index() & 63, nwaves = B >> 6;
  constexpr int NW = 8;  // qty, price, disc, cnt, dp.lo, dp.hi, ch.lo, ch.hi
  uint64_t* red = lds + shared_table_bytes(B) / 8;  // nwaves * 6 * NW words
#pragma unroll
  for (int g = 0; g < kNumGroups; g++) {
    uint64_t s64[4] = {qty[g], price[g], disc[g], cnt[g]};
    Wide128 w[2] = {dp[g], ch[g]};
    if (ballot(cnt[g] != 0)) {
#pragma unroll
      for (int f = 0; f < 4; f++) s64[f] = wave_sum_u64(s64[f]);
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) {
#pragma unroll
        for (int f = 0; f < 2; f++) {
          const uint64_t olo = __shfl_xor(w[f].lo, m, 64), ohi = __shfl_xor(w[f].hi, m, 64);
          const uint64_t nlo = w[f].lo + olo;
          w[f].hi += ohi + (nlo < olo);
          w[f].lo = nlo;
        }
      }
    } else {
      s64[0] = s64[1] = s64[2] = s64[3] = 0;
      w[0].lo = w[0].hi = w[1].lo = w[1].hi = 0;
    }
    if (lane == 0) {
      uint64_t* r = red + (wave * kNumGroups + g) * NW;
      r[0] = s64[0]; r[1] = s64[1]; r[2] = s64[2]; r[3] = s64[3];
      r[4] = w[0].lo; r[5] = w[0].hi; r[6] = w[1].lo; r[7] = w[1].hi;
    }
  }
  __syncthreads();
  if (thread_index() < kNumGroups) {
    const int g = thread_index();
    uint64_t s[4] = {0, 0, 0, 0};
    Wide128 w[2] = {{0, 0}, {0, 0}};
    for (uint32_t v = 0; v < nwaves; v++) {
      const uint64_t* r = red + (v * kNumGroups + g) * NW;
      for (int f = 0; f < 4; f++) s[f] += r[f];
      for (int f = 0; f < 2; f++) {
        const uint64_t nlo = w[f].lo + r[4 + 2 * f];
        w[f].hi += r[5 + 2 * f] + (nlo < w[f].lo);
        w[f].lo = nlo;
      }
    }
    if (s[3]) {
      const uint64_t qs =
          (L == Q1_LAYOUT_COMPRESSED || L == Q1_LAYOUT_COMPRESSED_PUSHDOWN) ? kQuantityScaleCompressed : 1;
      atomicAdd((unsigned long long*)&a.out->sum_qty[g], (unsigned long long)(s[0] * qs));
      atomicAdd((unsigned long long*)&a.out->sum_base_price[g], (unsigned long long)s[1]);
      atomicAdd((unsigned long long*)&a.out->sum_disc[g], (unsigned long long)s[2]);
      atomicAdd((unsigned long long*)&a.out->count[g], (unsigned long long)s[3]);
      atomic_add128(&a.out->sum_disc_price_lo[g], &a.out->sum_disc_price_hi[g], w[0].lo, w[0].hi);
      atomic_add128(&a.out->sum_charge_lo[g], &a.out->sum_charge_hi[g], w[1].lo, w[1].hi);
    }
  }
}

// The any-value kernel for tables known to hold values outside the domain (q1_run and the
// server on such tables): 16-row tiles, each lane's tile tested as above, and the wide
// sums in registers, so a table with every row outside the domain still runs at a few
// times the fast kernel's time instead of the private-memory table's.
template <int L, int B>
__global__ void __launch_bounds__(B)
scan_wide(ScanArgs a, uint32_t tiles_per_thread) {
  using Tile = typename TileOf<L>::T;
  constexpr int TILE = Tile::TILE;
  extern __shared__ uint64_t lds[];

  const int32_t cut = column_cutoff<L>(a.cutoff_day);
  const uint64_t first_tile = a.row_begin / kRowAlign * kRowAlign / TILE;  // as the host counts them
  const uint64_t end_tile = (a.row_end + TILE - 1) / TILE;
  const uint64_t base = first_tile + (uint64_t)blockIdx.x * B * tiles_per_thread;

  AccShared<B> narrow;
  narrow.init(reinterpret_cast<char*>(lds));
  uint64_t qty[kNumGroups], price[kNumGroups], disc[kNumGroups], cnt[kNumGroups];
  Wide128 dp[kNumGroups], ch[kNumGroups];
#pragma unroll
  for (int g = 0; g < kNumGroups; g++) {
    qty[g] = price[g] = disc[g] = cnt[g] = 0;
    dp[g].lo = dp[g].hi = ch[g].lo = ch[g].hi = 0;
  }

  constexpr int U = Tile::GROUP;
  for (uint32_t k = 0; k < tiles_per_thread; k += U) {
    TileGroup<Tile, U> grp;
    grp.load(a.cols, base + (uint64_t)k * B + thread_index(), B, tiles_per_thread - k, end_tile, a.row_begin,
             a.row_end);
#pragma unroll
    for (int m = 0; m < U; m++) {
      if (!grp.use[m]) continue;
      const Tile& tile = grp.tile[m];
      if (tile.in_domain()) {
        if (!grp.edge[m]) {
#pragma unroll
          for (int j = 0; j < TILE; j++) narrow.add(tile.template fast<false>(j, cut, 0));
        } else {
#pragma unroll
          for (int j = 0; j < TILE; j++) narrow.add(tile.template fast<true>(j, cut, grp.valid[m]));
        }
      } else {
#pragma unroll
        for (int j = 0; j < TILE; j++) {
          const WideRow r = tile.wide(j, cut, grp.valid[m]);
          const int64_t disc_1 = (int64_t)(100ull - (uint64_t)r.disc);
          const int64_t disc_price = (int64_t)((uint64_t)disc_1 * (uint64_t)r.price);
          const int64_t charge = (int64_t)((uint64_t)disc_price * (uint64_t)(r.tax + 100));
#pragma unroll
          for (int g = 0; g < kNumGroups; g++) {
            if (r.gid == (uint32_t)g) {
              qty[g] += (uint64_t)r.qty;
              price[g] += (uint64_t)r.price;
              disc[g] += (uint64_t)r.disc;
              cnt[g] += 1;
              dp[g].add64s(disc_price);
              ch[g].add64s(charge);
            }
          }
        }
      }
    }
  }

  // Fold the lane's narrow table into its wide one.  Lane sums of the narrow table are
  // exact and far below 2^63 (at most 16384 rows per lane).
  {
    LaneTotals nt;
    narrow.totals(nt, reinterpret_cast<char*>(lds));
#pragma unroll
    for (int g = 0; g < kNumGroups; g++) {
      qty[g] += nt.v[g][F_QTY];
      price[g] += nt.v[g][F_PRICE];
      disc[g] += nt.v[g][F_DISC];
      cnt[g] += nt.v[g][F_CNT];
      dp[g].add64s((int64_t)(100u * nt.v[g][F_PRICE] - nt.v[g][F_DPROD]));
      ch[g].add64s((int64_t)nt.v[g][F_CHARGE]);
    }
  }

  // Block merge in 128 bits for every field; the scratch follows the narrow table.
  const uint32_t wave = thread_index() >> 6, lane = thread_index() & 63, nwaves = B >> 6;
  constexpr int NW = 8;  // qty, price, disc, cnt, dp.lo, dp.hi, ch.lo, ch.hi
  uint64_t* red = lds + shared_table_bytes(B) / 8;  // nwaves * 6 * NW words
#pragma unroll
  for (int g = 0; g < kNumGroups; g++) {
    uint64_t s64[4] = {qty[g], price[g], disc[g], cnt[g]};
    Wide128 w[2] = {dp[g], ch[g]};
    if (ballot(cnt[g] != 0)) {
#pragma unroll
      for (int f = 0; f < 4; f++) s64[f] = wave_sum_u64(s64[f]);
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) {
#pragma unroll
        for (int f = 0; f < 2; f++) {
          const uint64_t olo = __shfl_xor(w[f].lo, m, 64), ohi = __shfl_xor(w[f].hi, m, 64);
          const uint64_t nlo = w[f].lo + olo;
          w[f].hi += ohi + (nlo < olo);
          w[f].lo = nlo;
        }
      }
    } else {
      s64[0] = s64[1] = s64[2] = s64[3] = 0;
      w[0].lo = w[0].hi = w[1].lo = w[1].hi = 0;
    }
    if (lane == 0) {
      uint64_t* r = red + (wave * kNumGroups + g) * NW;
      r[0] = s64[0]; r[1] = s64[1]; r[2] = s64[2]; r[3] = s64[3];
      r[4] = w[0].lo; r[5] = w[0].hi; r[6] = w[1].lo; r[7] = w[1].hi;
    }
  }
  __syncthreads();
  if (thread_index() < kNumGroups) {
    const int g = thread_index();
    uint64_t s[4] = {0, 0, 0, 0};
    Wide128 w[2] = {{0, 0}, {0, 0}};
    for (uint32_t v = 0; v < nwaves; v++) {
      const uint64_t* r = red + (v * kNumGroups + g) * NW;
      for (int f = 0; f < 4; f++) s[f] += r[f];
      for (int f = 0; f < 2; f++) {
        const uint64_t nlo = w[f].lo + r[4 + 2 * f];
        w[f].hi += r[5 + 2 * f] + (nlo < w[f].lo);
        w[f].lo = nlo;
      }
    }
    if (s[3]) {
      const uint64_t qs =
          (L == Q1_LAYOUT_COMPRESSED || L == Q1_LAYOUT_COMPRESSED_PUSHDOWN) ? kQuantityScaleCompressed : 1;
      atomicAdd((unsigned long long*)&a.out->sum_qty[g], (unsigned long long)(s[0] * qs));
      atomicAdd((unsigned long long*)&a.out->sum_base_price[g], (unsigned long long)s[1]);
      atomicAdd((unsigned long long*)&a.out->sum_disc[g], (unsigned long long)s[2]);
      atomicAdd((unsigned long long*)&a.out->count[g], (unsigned long long)s[3]);
      atomic_add128(&a.out->sum_disc_price_lo[g], &a.out->sum_disc_price_hi[g], w[0].lo, w[0].hi);
      atomic_add128(&a.out->sum_charge_lo[g], &a.out->sum_charge_hi[g], w[1].lo, w[1].hi);
    }
  }
}

// ------------------------------------------------------------------------------------
// The batched persistent kernel.
// ------------------------------------------------------------------------------------
// Column pointers read from a descriptor.  Pointers that arrive as kernel arguments are
// known to be global; generic ones loaded from memory make flat loads: 64-bit vector
// addresses, and flat loads count against the LDS wait counter too, so every wait for an
// LDS atomic would also wait for them.  Reading the pointers as global-qualified values
// keeps the tile loads global (a scalar base plus a vector offset).
typedef const __attribute__((address_space(1))) void* GlobalPtr;
struct GlobalColumns {
  GlobalPtr ship_date, quantity, extended_price, discount, tax, return_flag, line_status, filter;
};
static_assert(sizeof(GlobalColumns) == sizeof(q1_columns), "same layout as q1_columns");

__device__ __forceinline__ q1_columns load_global_columns(const q1_columns* src) {
  const GlobalColumns g = *reinterpret_cast<const GlobalColumns*>(src);
  q1_columns c;
  c.ship_date = (const void*)g.ship_date;
  c.quantity = (const void*)g.quantity;
  c.extended_price = (const void*)g.extended_price;
  c.discount = (const void*)g.discount;
  c.tax = (const void*)g.tax;
  c.return_flag = (const void*)g.return_flag;
  c.line_status = (const void*)g.line_status;
  c.filter = (const void*)g.filter;
  return c;
}

// The last request whose first block is <= v, searching forward from request r (whose
// first block is <= v).  First blocks never decrease along the list, so the lanes that
// pass form a prefix of the wavefront; every wavefront of the block computes the same r.
// next_begin receives the first block of the request after it (total past the last), so
// later entries below it need no search.
__device__ __forceinline__ uint32_t find_request(const BatchDesc* descs, uint32_t n, uint32_t r, uint64_t v,
                                                 uint64_t total, uint64_t& next_begin) {
  const uint32_t lane = thread_index() & 63;
  for (;;) {
    const uint32_t i = r + 1 + lane;
    const uint64_t b = i < n ? descs[i].block_begin : total;  // total > v
    const uint32_t pass = (uint32_t)__popcll(ballot(b <= v));
    r += pass;
    if (pass < 64) {
      next_begin = (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, (int)pass) |
                   (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), (int)pass) << 32;
      return r;
    }
  }
}

// The batched kernel: the requests' scan blocks (request r's block k is the block k that
// launch_scan would run over r's rows alone, with hdr->tiles_per_thread) form one list of
// hdr->total_blocks blocks.  A resident grid works through the list with a fixed stride:
// block b runs entries b, b + gridDim.x, b + 2 gridDim.x, ...  The blocks in flight
// therefore read neighbouring stretches of HBM whatever the request sizes, with no shared
// counter (a claim counter that every block updates once per entry cost 6-11 % of the read
// rate of this structure).  Each entry runs the scan kernel's block body and merges into
// one of its request's 64 spread tables (entry v into table v mod 64) when the request has
// them, else into its request's result, so that the blocks of one request, which run in
// step, do not queue on one table's atomics.  Every block exits once its next entry lies
// past the list.
// The header and descriptors come as read-only pointers, so they load through the scalar
// cache into scalar registers.
template <int L, int P>
__global__ void __launch_bounds__(kBatchBlock)
batch_scan(const BatchHeader* __restrict__ hdr, const BatchDesc* __restrict__ descs, q1_device_aggregates* outs,
           q1_device_aggregates* spread) {
  extern __shared__ uint64_t lds[];
  const uint32_t n = (uint32_t)hdr->n, tpt = hdr->tiles_per_thread, spread_requests = hdr->spread_requests;
  const uint64_t total = hdr->total_blocks;
  uint64_t v = blockIdx.x, next_begin = 0;
  uint32_t r = 0;
  while (v < total) {
    if (v >= next_begin) r = find_request(descs, n, r, v, total, next_begin);
    const BatchDesc d = descs[r];
    q1_device_aggregates* out =
        r < spread_requests ? spread + (uint64_t)r * kSpreadTables + (v & (kSpreadTables - 1)) : outs + r;
    ScanArgs a;
    a.cols = load_global_columns(&descs[r].cols);
    a.row_begin = d.row_begin;
    a.row_end = d.row_end;
    a.cutoff_day = d.cutoff_day;
    a.out = out;
    a.out_mask = 0;
    scan_block<L, P, kBatchBlock>(a, v - d.block_begin, tpt, nullptr, lds, out);
    // Orders this entry's reads of the merge scratch before the next entry's writes.
    __syncthreads();
    v += gridDim.x;
  }
}

// ------------------------------------------------------------------------------------
// Dispatch templates: the kernel of a (layout, placement, block size).
// ------------------------------------------------------------------------------------
template <int L, int B>
constexpr FastFn fast_fn_b(int p) {
  return p == Q1_PLACE_IN_REGISTERS ? scan_fast<L, Q1_PLACE_IN_REGISTERS, B>
       : p == Q1_PLACE_IN_REGISTERS_PER_THREAD ? scan_fast<L, Q1_PLACE_IN_REGISTERS_PER_THREAD, B>
       : p == Q1_PLACE_LOCAL_MEM ? scan_fast<L, Q1_PLACE_LOCAL_MEM, B>
       : p == Q1_PLACE_PER_THREAD_SHARED_MEM ? scan_fast<L, Q1_PLACE_PER_THREAD_SHARED_MEM, B>
       : p == kPlaceSharedNarrow ? scan_fast<L, kPlaceSharedNarrow, B>
       : p == kPlaceSharedPacked ? scan_fast<L, kPlaceSharedPacked, B>
       : scan_fast<L, Q1_PLACE_GLOBAL, B>;
}

template <int L>
FastFn fast_fn(int p, uint32_t threads) {
  return threads == 64 ? fast_fn_b<L, 64>(p) : threads == 128 ? fast_fn_b<L, 128>(p) : fast_fn_b<L, 256>(p);
}

template <int L>
GenFn any_fn(uint32_t threads) {
  return threads == 64 ? scan_any<L, 64> : threads == 128 ? scan_any<L, 128> : scan_any<L, 256>;
}

template <int L>
GenFn wide_fn(uint32_t threads) {
  return threads == 64 ? scan_wide<L, 64> : threads == 128 ? scan_wide<L, 128> : scan_wide<L, 256>;
}

// The batched kernel's variants: the shared placement runs the narrow per-lane table (the
// packed one for the compressed layouts), as launch_scan does for at most 256 rows per lane.
template <int L>
void* batch_fn(q1_placement p) {
  constexpr bool units = L == Q1_LAYOUT_COMPRESSED || L == Q1_LAYOUT_COMPRESSED_PUSHDOWN;
  if (p != Q1_PLACE_PER_THREAD_SHARED_MEM) return (void*)batch_scan<L, Q1_PLACE_IN_REGISTERS>;
  return units ? (void*)batch_scan<L, kPlaceSharedPacked> : (void*)batch_scan<L, kPlaceSharedNarrow>;
}

}  // namespace

// Kernel entry points of one layout, defined by that layout's translation unit.
#define Q1_DECLARE_LAYOUT_KERNELS(NAME)                  \
  FastFn fast_kernel_##NAME(int placement, uint32_t threads); \
  GenFn any_kernel_##NAME(uint32_t threads);             \
  GenFn wide_kernel_##NAME(uint32_t threads);            \
  void* batch_kernel_##NAME(q1_placement placement);
Q1_DECLARE_LAYOUT_KERNELS(reference)
Q1_DECLARE_LAYOUT_KERNELS(uncompressed)
Q1_DECLARE_LAYOUT_KERNELS(compressed)
Q1_DECLARE_LAYOUT_KERNELS(pushdown)

#define Q1_DEFINE_LAYOUT_KERNELS(L, NAME)                                                          \
  FastFn fast_kernel_##NAME(int placement, uint32_t threads) { return fast_fn<L>(placement, threads); } \
  GenFn any_kernel_##NAME(uint32_t threads) { return any_fn<L>(threads); }                         \
  GenFn wide_kernel_##NAME(uint32_t threads) { return wide_fn<L>(threads); }                       \
  void* batch_kernel_##NAME(q1_placement placement) { return batch_fn<L>(placement); }

}  // namespace q1
