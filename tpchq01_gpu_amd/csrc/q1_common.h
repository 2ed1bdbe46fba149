// Shared host/device definitions for the TPC-H Q1 scan -> filter -> group-aggregate path:
// constants, dictionary encodings, the counter-based synthetic lineitem generator and
// calendar arithmetic.  Everything here is plain integer code usable from both the
// host compiler and the gfx950 device pass.
#pragma once

#include <stdint.h>

#if defined(__HIP__)
#define Q1_HD __host__ __device__ __forceinline__
#else
#define Q1_HD static inline
#endif

namespace q1 {

// ---------------------------------------------------------------------------------
// Constants (src/constants.hpp:166-186 of the reference)
// ---------------------------------------------------------------------------------
enum : int32_t {
  kShipDateFrameOfReference = 727563,  // 1992-01-01 as a MonetDB day number
  kDefaultCutoff = 729999,             // 1998-12-01 minus 90 days = 1998-09-02
  kNumGroups = 6,                      // 3 return flags x 2 line statuses, slot = rf<<1 | ls
  kRowsPerLaneTile = 16,               // rows one lane handles per tile (one 16-byte u8 load)
  kRfPerWord = 16,                     // 2-bit return flags per 32-bit word
  kLsPerWord = 32,                     // 1-bit line statuses / filter bits per 32-bit word
};

// Scale of the quantity column per layout: the compressed layout stores whole units,
// the others keep Decimal(15,2) x100 (src/main.cu:783 narrows qty/100).
enum : int64_t { kQuantityScaleCompressed = 100 };

// ---------------------------------------------------------------------------------
// Dictionary encodings (src/data_types.h:52-80, release form):
//   return flag A->0, N->1, R->2 ; line status F->0, O->1 ; slot = rf<<1 | ls
// ---------------------------------------------------------------------------------
Q1_HD uint32_t encode_return_flag(uint32_t c) { return ((c == 'R') << 1) | (c == 'N'); }
Q1_HD uint32_t encode_line_status(uint32_t c) { return c == 'F' ? 0u : 1u; }
Q1_HD char decode_return_flag(uint32_t e) { return e == 0 ? 'A' : (e == 1 ? 'N' : 'R'); }
Q1_HD char decode_line_status(uint32_t e) { return e == 0 ? 'F' : 'O'; }

// ---------------------------------------------------------------------------------
// Calendar: MonetDB day numbers (days since 0000-01-01 counted the way cpu/monetdb.cpp
// todate() does) equal the proleptic Gregorian day ordinal (0001-01-01 = 1) plus 365.
// days_from_civil below is the standard era-based civil-to-days algorithm.
// ---------------------------------------------------------------------------------
Q1_HD int32_t day_number(int32_t year, int32_t month, int32_t day) {
  const int32_t y = year - (month <= 2);
  const int32_t era = (y >= 0 ? y : y - 399) / 400;
  const int32_t yoe = y - era * 400;                                        // [0, 399]
  const int32_t doy = (153 * (month + (month > 2 ? -3 : 9)) + 2) / 5 + day - 1;  // [0, 365]
  const int32_t doe = yoe * 365 + yoe / 4 - yoe / 100 + doy;              // [0, 146096]
  const int32_t days_since_1970 = era * 146097 + doe - 719468;
  // 1970-01-01 has Gregorian ordinal 719163, hence MonetDB day 719528.
  return days_since_1970 + 719528;
}

// ---------------------------------------------------------------------------------
// Synthetic lineitem generator.  Counter-based: every value is a pure function of
// (seed, column stream, row), so any row range regenerates identically on the host
// and on the device, and a table of any size can be produced in parallel.
//   shipdate  uniform day in 1992-01-02 .. 1998-12-01
//   receipt   shipdate + uniform 1..30
//   returnflag R or A (50/50) if receipt <= 1995-06-17, else N
//   linestatus F if shipdate <= 1995-06-17, else O
//   quantity  uniform 1..50, part price uniform 900.00..2099.99 (cents),
//   extendedprice = quantity x part price, discount 0..10, tax 0..8 (hundredths)
// ---------------------------------------------------------------------------------
enum : uint32_t {
  kStreamShip = 1, kStreamReceipt = 2, kStreamFlag = 3, kStreamQty = 4,
  kStreamPrice = 5, kStreamDisc = 6, kStreamTax = 7,
};
enum : int32_t {
  kGenFirstShip = 727564,   // 1992-01-02
  kGenShipDays = 2526,      // through 1998-12-01 inclusive
  kGenStatusCut = 728826,   // 1995-06-17
};

Q1_HD uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Uniform integer in [0, range) from the high 32 bits of a keyed hash.
Q1_HD uint32_t gen_uniform(uint64_t seed, uint32_t stream, uint64_t row, uint32_t range) {
  const uint64_t h = mix64(seed ^ mix64((row << 3) | stream));
  return (uint32_t)(((h >> 32) * (uint64_t)range) >> 32);
}

struct GenRow {
  int32_t ship_date;        // MonetDB day number
  int32_t quantity;         // whole units 1..50
  int64_t extended_price;   // cents
  int32_t discount;         // hundredths 0..10
  int32_t tax;              // hundredths 0..8
  char return_flag;         // 'A' | 'N' | 'R'
  char line_status;         // 'F' | 'O'
};

Q1_HD GenRow gen_row(uint64_t seed, uint64_t row) {
  GenRow r;
  r.ship_date = kGenFirstShip + (int32_t)gen_uniform(seed, kStreamShip, row, kGenShipDays);
  const int32_t receipt = r.ship_date + 1 + (int32_t)gen_uniform(seed, kStreamReceipt, row, 30);
  r.return_flag = receipt <= kGenStatusCut
                      ? (gen_uniform(seed, kStreamFlag, row, 2) ? 'R' : 'A')
                      : 'N';
  r.line_status = r.ship_date <= kGenStatusCut ? 'F' : 'O';
  r.quantity = 1 + (int32_t)gen_uniform(seed, kStreamQty, row, 50);
  const int64_t part_price = 90000 + (int64_t)gen_uniform(seed, kStreamPrice, row, 120000);
  r.extended_price = (int64_t)r.quantity * part_price;
  r.discount = (int32_t)gen_uniform(seed, kStreamDisc, row, 11);
  r.tax = (int32_t)gen_uniform(seed, kStreamTax, row, 9);
  return r;
}

}  // namespace q1
