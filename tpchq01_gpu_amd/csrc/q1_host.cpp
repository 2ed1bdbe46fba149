// Host-side parts of the path: the seeded generator for host columns, the column
// compression of src/main.cu:773-803, the pushdown producer of src/main.cu:362-389 and the
// CPU coprocessor (morsel-parallel aggregation in the manner of cpu/kernel_x100.hpp:456-500:
// workers claim fixed-size morsels with an atomic counter and merge private tables).
#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/q1_hip.h"
#include "q1_common.h"
#include "q1_kernels.h"

namespace {

using i128 = __int128;

int thread_count(int requested) {
  if (requested > 0) return requested;
  const unsigned hc = std::thread::hardware_concurrency();
  return hc ? (int)hc : 1;
}

// Runs fn(begin, end) over [0, n) split into `threads` pieces aligned to `align` items.
template <class F>
void parallel_ranges(uint64_t n, uint64_t align, int threads, F&& fn) {
  if (n == 0) return;
  threads = thread_count(threads);
  uint64_t per = (n + threads - 1) / threads;
  per = (per + align - 1) / align * align;
  std::vector<std::thread> ts;
  for (uint64_t b = 0; b < n; b += per) {
    const uint64_t e = b + per < n ? b + per : n;
    ts.emplace_back([&fn, b, e] { fn(b, e); });
  }
  for (auto& t : ts) t.join();
}

constexpr bool compressed_family(q1_layout l) { return l == Q1_LAYOUT_COMPRESSED || l == Q1_LAYOUT_COMPRESSED_PUSHDOWN; }

// ---- row access per layout --------------------------------------------------------------
struct RowVals {
  int64_t ship, qty, price, disc, tax;
  uint32_t slot;  // 0..5, or >= 6 for keys outside A/N/R x F/O (skipped, as the kernels do)
};

template <int L>
inline RowVals read_row(const q1_columns& c, uint64_t i) {
  RowVals r;
  if (L == Q1_LAYOUT_REFERENCE) {
    r.ship = ((const int32_t*)c.ship_date)[i];
    r.qty = ((const int64_t*)c.quantity)[i];
    r.price = ((const int64_t*)c.extended_price)[i];
    r.disc = ((const int64_t*)c.discount)[i];
    r.tax = ((const int64_t*)c.tax)[i];
    r.slot = q1::encode_return_flag((uint8_t)((const char*)c.return_flag)[i]) << 1 |
             q1::encode_line_status((uint8_t)((const char*)c.line_status)[i]);
  } else if (L == Q1_LAYOUT_UNCOMPRESSED) {
    r.ship = ((const int32_t*)c.ship_date)[i];
    r.qty = ((const int32_t*)c.quantity)[i];
    r.price = ((const int64_t*)c.extended_price)[i];
    r.disc = ((const int8_t*)c.discount)[i];
    r.tax = ((const int8_t*)c.tax)[i];
    r.slot = q1::encode_return_flag((uint8_t)((const char*)c.return_flag)[i]) << 1 |
             q1::encode_line_status((uint8_t)((const char*)c.line_status)[i]);
  } else {
    r.ship = L == Q1_LAYOUT_COMPRESSED ? ((const uint16_t*)c.ship_date)[i] : 0;
    r.qty = ((const uint8_t*)c.quantity)[i];
    r.price = ((const uint32_t*)c.extended_price)[i];
    r.disc = ((const uint8_t*)c.discount)[i];
    r.tax = ((const uint8_t*)c.tax)[i];
    const uint32_t rf = (((const uint32_t*)c.return_flag)[i / 16] >> (2 * (i % 16))) & 3u;
    const uint32_t ls = (((const uint32_t*)c.line_status)[i / 32] >> (i % 32)) & 1u;
    r.slot = rf << 1 | ls;
  }
  return r;
}

// One worker's private table; cache-line aligned so that neighbouring workers' tables
// share no line.
struct alignas(64) Table6 {
  uint64_t qty[6], price[6], disc[6], cnt[6];
  i128 dp[6], ch[6];
  void clear() { std::memset(this, 0, sizeof(*this)); }
  void add(const Table6& o) {
    for (int g = 0; g < 6; g++) {
      qty[g] += o.qty[g]; price[g] += o.price[g]; disc[g] += o.disc[g]; cnt[g] += o.cnt[g];
      dp[g] += o.dp[g]; ch[g] += o.ch[g];
    }
  }
};

// Compressed columns hold u8 discount and tax and a u32 price, so a row's disc_price is
// below 2^40 and its charge below 2^48 in magnitude, and 2^14 rows of them sum exactly in
// int64: those layouts add the two products in int64 per run of kRun rows and widen to
// 128 bits once per run.  The wide layouts' int64 columns keep KernelNaive's wrapping
// int64 products with 128-bit sums per row.
constexpr uint64_t kRun = 1u << 14;

template <int L>
void aggregate_range(const q1_columns& c, uint64_t b, uint64_t e, int64_t cut, Table6& out) {
  if (compressed_family((q1_layout)L) && e - b > kRun) {
    for (uint64_t r = b; r < e; r += kRun) aggregate_range<L>(c, r, r + kRun < e ? r + kRun : e, cut, out);
    return;
  }
  // A table local to the morsel: the column pointers (char / uint8 columns may alias any
  // object) cannot point into it, so its sums stay out of the loads and stores of the
  // worker's shared table until the morsel ends.
  Table6 t;
  t.clear();
  int64_t dp64[6] = {}, ch64[6] = {};
  for (uint64_t i = b; i < e; i++) {
    bool pass;
    if (L == Q1_LAYOUT_COMPRESSED_PUSHDOWN)
      pass = (((const uint32_t*)c.filter)[i / 32] >> (i % 32)) & 1u;
    else
      pass = ((L == Q1_LAYOUT_COMPRESSED) ? (int64_t)((const uint16_t*)c.ship_date)[i]
                                          : (int64_t)read_row<L>(c, i).ship) <= cut;
    if (!pass) continue;
    const RowVals r = read_row<L>(c, i);
    if (r.slot >= 6) continue;
    const int64_t disc_1 = (int64_t)(100ull - (uint64_t)r.disc);
    const int64_t disc_price = (int64_t)((uint64_t)disc_1 * (uint64_t)r.price);
    const int64_t charge = (int64_t)((uint64_t)disc_price * (uint64_t)(r.tax + 100));
    const uint32_t g = r.slot;
    t.qty[g] += (uint64_t)r.qty;
    t.price[g] += (uint64_t)r.price;
    t.disc[g] += (uint64_t)r.disc;
    t.cnt[g] += 1;
    if (compressed_family((q1_layout)L)) {
      dp64[g] += disc_price;
      ch64[g] += charge;
    } else {
      t.dp[g] += (i128)disc_price;
      t.ch[g] += (i128)charge;
    }
  }
  if (compressed_family((q1_layout)L)) {
    for (int g = 0; g < 6; g++) {
      t.dp[g] += (i128)dp64[g];
      t.ch[g] += (i128)ch64[g];
    }
  }
  out.add(t);
}

using RangeFn = void (*)(const q1_columns&, uint64_t, uint64_t, int64_t, Table6&);

RangeFn range_fn(q1_layout l) {
  switch (l) {
    case Q1_LAYOUT_REFERENCE: return aggregate_range<Q1_LAYOUT_REFERENCE>;
    case Q1_LAYOUT_UNCOMPRESSED: return aggregate_range<Q1_LAYOUT_UNCOMPRESSED>;
    case Q1_LAYOUT_COMPRESSED: return aggregate_range<Q1_LAYOUT_COMPRESSED>;
    default: return aggregate_range<Q1_LAYOUT_COMPRESSED_PUSHDOWN>;
  }
}

}  // namespace

extern "C" {

q1_status q1_generate_host(q1_layout layout, uint64_t seed, uint64_t row_begin, uint64_t rows,
                           const q1_columns* out, int threads) {
  if (!out || layout < Q1_LAYOUT_REFERENCE || layout > Q1_LAYOUT_COMPRESSED) return Q1_EINVAL;
  if (layout == Q1_LAYOUT_COMPRESSED && row_begin % 32) return Q1_EINVAL;
  const q1_columns c = *out;
  if (!c.ship_date || !c.quantity || !c.extended_price || !c.discount || !c.tax || !c.return_flag ||
      !c.line_status)
    return Q1_EINVAL;
  try {
    parallel_ranges(rows, 32, threads, [&](uint64_t b, uint64_t e) {
      // Each range starts on a 32-row boundary, so its packed words are its own.
      uint32_t rfw = 0, lsw = 0;
      for (uint64_t i = b; i < e; i++) {
        const q1::GenRow g = q1::gen_row(seed, row_begin + i);
        if (layout == Q1_LAYOUT_REFERENCE) {
          ((int32_t*)c.ship_date)[i] = g.ship_date;
          ((int64_t*)c.quantity)[i] = (int64_t)g.quantity * 100;
          ((int64_t*)c.extended_price)[i] = g.extended_price;
          ((int64_t*)c.discount)[i] = g.discount;
          ((int64_t*)c.tax)[i] = g.tax;
          ((char*)c.return_flag)[i] = g.return_flag;
          ((char*)c.line_status)[i] = g.line_status;
        } else if (layout == Q1_LAYOUT_UNCOMPRESSED) {
          ((int32_t*)c.ship_date)[i] = g.ship_date;
          ((int32_t*)c.quantity)[i] = g.quantity * 100;
          ((int64_t*)c.extended_price)[i] = g.extended_price;
          ((int8_t*)c.discount)[i] = (int8_t)g.discount;
          ((int8_t*)c.tax)[i] = (int8_t)g.tax;
          ((char*)c.return_flag)[i] = g.return_flag;
          ((char*)c.line_status)[i] = g.line_status;
        } else {
          ((uint16_t*)c.ship_date)[i] = (uint16_t)(g.ship_date - q1::kShipDateFrameOfReference);
          ((uint8_t*)c.quantity)[i] = (uint8_t)g.quantity;
          ((uint32_t*)c.extended_price)[i] = (uint32_t)g.extended_price;
          ((uint8_t*)c.discount)[i] = (uint8_t)g.discount;
          ((uint8_t*)c.tax)[i] = (uint8_t)g.tax;
          rfw |= q1::encode_return_flag((uint32_t)g.return_flag) << (2 * (i % 16));
          lsw |= q1::encode_line_status((uint32_t)g.line_status) << (i % 32);
          if (i % 16 == 15 || i + 1 == e) {
            ((uint32_t*)c.return_flag)[i / 16] = rfw;
            rfw = 0;
          }
          if (i % 32 == 31 || i + 1 == e) {
            ((uint32_t*)c.line_status)[i / 32] = lsw;
            lsw = 0;
          }
        }
      }
    });
  } catch (...) {
    return Q1_ENOMEM;
  }
  return Q1_OK;
}

q1_status q1_compress_columns(const q1_columns* in, uint64_t rows, const q1_columns* out, int threads) {
  if (!in || !out) return Q1_EINVAL;
  const q1_columns a = *in, c = *out;
  try {
    parallel_ranges(rows, 32, threads, [&](uint64_t b, uint64_t e) {
      for (uint64_t i = b; i < e; i++) {
        ((uint16_t*)c.ship_date)[i] = (uint16_t)(((const int32_t*)a.ship_date)[i] - q1::kShipDateFrameOfReference);
        ((uint8_t*)c.discount)[i] = (uint8_t)((const int64_t*)a.discount)[i];
        ((uint32_t*)c.extended_price)[i] = (uint32_t)((const int64_t*)a.extended_price)[i];
        ((uint8_t*)c.quantity)[i] = (uint8_t)(((const int64_t*)a.quantity)[i] / 100);
        ((uint8_t*)c.tax)[i] = (uint8_t)((const int64_t*)a.tax)[i];
      }
      for (uint64_t w = b / 16; w < (e + 15) / 16; w++) {
        uint32_t word = 0;
        for (uint64_t i = w * 16; i < w * 16 + 16 && i < e; i++)
          word |= q1::encode_return_flag((uint8_t)((const char*)a.return_flag)[i]) << (2 * (i % 16));
        ((uint32_t*)c.return_flag)[w] = word;
      }
      for (uint64_t w = b / 32; w < (e + 31) / 32; w++) {
        uint32_t word = 0;
        for (uint64_t i = w * 32; i < w * 32 + 32 && i < e; i++)
          word |= q1::encode_line_status((uint8_t)((const char*)a.line_status)[i]) << (i % 32);
        ((uint32_t*)c.line_status)[w] = word;
      }
    });
  } catch (...) {
    return Q1_ENOMEM;
  }
  return Q1_OK;
}

q1_status q1_cpu_build_filter(const uint16_t* ship, uint64_t rows, int32_t cutoff_day, uint32_t* out,
                              int threads) {
  if ((!ship || !out) && rows) return Q1_EINVAL;
  const int64_t cut = (int64_t)cutoff_day - q1::kShipDateFrameOfReference;
  try {
    parallel_ranges((rows + 31) / 32, 1, threads, [&](uint64_t wb, uint64_t we) {
      for (uint64_t w = wb; w < we; w++) {
        uint32_t word = 0;
        for (uint64_t i = w * 32; i < w * 32 + 32 && i < rows; i++) word |= (uint32_t)((int64_t)ship[i] <= cut) << (i % 32);
        out[w] = word;
      }
    });
  } catch (...) {
    return Q1_ENOMEM;
  }
  return Q1_OK;
}

q1_status q1_host_columns_in_tpch_domain(q1_layout layout, const q1_columns* host, uint64_t rows, int threads,
                                         int* in_domain) {
  if (!host || !in_domain || (int)layout < 0 || layout > Q1_LAYOUT_COMPRESSED_PUSHDOWN) return Q1_EINVAL;
  if (rows && (!host->quantity || !host->extended_price || !host->discount || !host->tax)) return Q1_EINVAL;
  struct MinMax {
    int64_t lo[4] = {INT64_MAX, INT64_MAX, INT64_MAX, INT64_MAX};
    int64_t hi[4] = {INT64_MIN, INT64_MIN, INT64_MIN, INT64_MIN};
  };
  std::vector<MinMax> parts(thread_count(threads) + 1);
  std::atomic<int> next{0};
  try {
    parallel_ranges(rows, 1, threads, [&](uint64_t b, uint64_t e) {
      MinMax m;
      for (uint64_t i = b; i < e; i++) {
        int64_t v[4];
        if (layout == Q1_LAYOUT_REFERENCE) {
          v[0] = ((const int64_t*)host->discount)[i];
          v[1] = ((const int64_t*)host->tax)[i];
          v[2] = ((const int64_t*)host->quantity)[i];
          v[3] = ((const int64_t*)host->extended_price)[i];
        } else if (layout == Q1_LAYOUT_UNCOMPRESSED) {
          v[0] = ((const int8_t*)host->discount)[i];
          v[1] = ((const int8_t*)host->tax)[i];
          v[2] = ((const int32_t*)host->quantity)[i];
          v[3] = ((const int64_t*)host->extended_price)[i];
        } else {
          v[0] = ((const uint8_t*)host->discount)[i];
          v[1] = ((const uint8_t*)host->tax)[i];
          v[2] = ((const uint8_t*)host->quantity)[i];
          v[3] = ((const uint32_t*)host->extended_price)[i];
        }
        for (int k = 0; k < 4; k++) {
          m.lo[k] = v[k] < m.lo[k] ? v[k] : m.lo[k];
          m.hi[k] = v[k] > m.hi[k] ? v[k] : m.hi[k];
        }
      }
      parts[next++] = m;
    });
  } catch (...) {
    return Q1_ENOMEM;
  }
  MinMax all;
  for (const MinMax& m : parts)
    for (int k = 0; k < 4; k++) {
      all.lo[k] = m.lo[k] < all.lo[k] ? m.lo[k] : all.lo[k];
      all.hi[k] = m.hi[k] > all.hi[k] ? m.hi[k] : all.hi[k];
    }
  const q1::Domain d{all.lo[0], all.hi[0], all.lo[1], all.hi[1], all.lo[2], all.hi[2], all.lo[3], all.hi[3]};
  *in_domain = rows == 0 || q1::domain_is_tpch(layout, d);
  return Q1_OK;
}

q1_status q1_cpu_run(q1_layout layout, const q1_columns* host, uint64_t row_begin, uint64_t row_count,
                     int32_t cutoff_day, int threads, q1_result* out) {
  if (!host || !out || layout < Q1_LAYOUT_REFERENCE || layout > Q1_LAYOUT_COMPRESSED_PUSHDOWN) return Q1_EINVAL;
  const int64_t cut = compressed_family(layout) ? (int64_t)cutoff_day - q1::kShipDateFrameOfReference : cutoff_day;
  const RangeFn fn = range_fn(layout);
  const uint64_t end = row_begin + row_count;
  const uint64_t morsel = 16 * 1024;
  threads = thread_count(threads);
  std::atomic<uint64_t> next{row_begin};
  std::vector<Table6> tabs((size_t)threads);
  try {
    std::vector<std::thread> ws;
    for (int t = 0; t < threads; t++) {
      ws.emplace_back([&, t] {
        Table6& tab = tabs[(size_t)t];
        tab.clear();
        for (;;) {
          const uint64_t off = next.fetch_add(morsel);
          if (off >= end) break;
          fn(*host, off, off + morsel < end ? off + morsel : end, cut, tab);
        }
      });
    }
    for (auto& w : ws) w.join();
  } catch (...) {
    return Q1_ENOMEM;
  }
  Table6 total;
  total.clear();
  for (const Table6& t : tabs) total.add(t);
  const uint64_t qs = compressed_family(layout) ? 100 : 1;
  for (int g = 0; g < 6; g++) {
    q1_group& r = out->g[g];
    r.sum_qty = (int64_t)(total.qty[g] * qs);
    r.sum_base_price = (int64_t)total.price[g];
    r.sum_disc = (int64_t)total.disc[g];
    r.count = (int64_t)total.cnt[g];
    r.sum_disc_price_lo = (uint64_t)total.dp[g];
    r.sum_disc_price_hi = (int64_t)(total.dp[g] >> 64);
    r.sum_charge_lo = (uint64_t)total.ch[g];
    r.sum_charge_hi = (int64_t)(total.ch[g] >> 64);
  }
  return Q1_OK;
}

}  // extern "C"
