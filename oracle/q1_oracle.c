/*
 * q1_oracle.c -- TEST INFRASTRUCTURE ONLY: the CPU restatement of the reference's TPC-H
 * Q1 algorithm that the HIP path is checked against (see q1_oracle.h).
 *
 * Build: oracle/Makefile (gcc, -O2).  Building with -DQ1O_MUTANT=<k> plants one known
 * mistake (numbered below); tests/test_oracle.py builds every mutant and requires that
 * at least one of the pinning checks rejects it, which shows the checks have teeth.
 *   1 discount and tax swapped in the charge      2 tax factor dropped from the charge
 *   3 discount added instead of subtracted        4 strict filter (<) instead of <=
 *   5 return flag and line status swapped in the slot
 *   6 return-flag bit offset off by one           7 line-status bit index taken mod 16
 *   8 charge accumulated in 64 bits               9 quantity not narrowed by 100
 *  10 last filter word dropped                    11 sum_disc accumulates (1 - disc)
 *  12 X100: the last partial vector of a morsel is dropped
 *  13 AVX-512: the last of the 8 table lanes is left out of the final sum
 *  14 compressed morsel-parallel: a range's last partial morsel is dropped
 */
#include "q1_oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#ifndef Q1O_MUTANT
#define Q1O_MUTANT 0
#endif

typedef __int128 i128;

/* ---- todate: cpu/monetdb.cpp:26-77 ------------------------------------------------ */
static const int kCumDays[13] = {0, 31, 59, 90, 120, 151, 181, 212, 243, 273, 304, 334, 365};

static int is_leap(int y) { return (y % 4 == 0) && ((y % 100 != 0) || (y % 400 == 0)); }

/* Leap years before Jan 1 of `year`+1 counted from year 0, as leapyears() (monetdb.cpp:47-59). */
static int leap_count(int year) { return year / 4 + year / 400 - year / 100 + (year >= 0); }

static int month_days(int m, int y) {
  static const int kDays[13] = {0, 31, 29, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
  return m != 2 ? kDays[m] : (is_leap(y) ? 29 : 28);
}

int32_t q1o_todate(int day, int month, int year) {
  /* validity check DATE(d,m,y) (monetdb.cpp:39); invalid dates give INT_MIN (date_nil) */
  if (!(month > 0 && month <= 12 && day > 0 && year != 0 && day <= month_days(month, year)))
    return (int32_t)0x80000000;
  if (year < 0) year++; /* hide year 0 (monetdb.cpp:67-68) */
  int32_t n = day - 1;
  if (month > 2 && is_leap(year)) n++;
  n += kCumDays[month - 1];
  n += 365 * year + leap_count(year >= 0 ? year - 1 : year);
  return n;
}

/* ---- shared aggregation step ---------------------------------------------------------- */
/* 6-slot index of a passing row's group, or -1 for keys outside {A,N,R} x {F,O}.  The
 * reference CPU kernels key rf<<8|ls into a 65535-slot table (cpu/kernel_naive.hpp:20);
 * the GPU kernels key encode_return_flag<<1 | encode_line_status (src/data_types.h:52-80).
 * For valid keys the two coincide up to this relabelling. */
static int slot_of(char rf, char ls) {
  int r, l;
  if (rf == 'A') r = 0; else if (rf == 'N') r = 1; else if (rf == 'R') r = 2; else return -1;
  if (ls == 'F') l = 0; else if (ls == 'O') l = 1; else return -1;
#if Q1O_MUTANT == 5
  return l * 3 + r;
#else
  return r << 1 | l;
#endif
}

typedef struct {
  int64_t sum_qty, sum_base_price, sum_disc, count;
  i128 sum_disc_price, sum_charge;
} acc_t;

static void acc_to_result(const acc_t* a, int64_t unexpected, q1o_result* out) {
  memset(out, 0, sizeof(*out));
  for (int g = 0; g < 6; g++) {
    out->g[g].sum_qty = a[g].sum_qty;
    out->g[g].sum_base_price = a[g].sum_base_price;
    out->g[g].sum_disc_price_lo = (uint64_t)a[g].sum_disc_price;
    out->g[g].sum_disc_price_hi = (int64_t)(a[g].sum_disc_price >> 64);
    out->g[g].sum_charge_lo = (uint64_t)a[g].sum_charge;
    out->g[g].sum_charge_hi = (int64_t)(a[g].sum_charge >> 64);
    out->g[g].sum_disc = a[g].sum_disc;
    out->g[g].count = a[g].count;
  }
  out->unexpected_keys = unexpected;
}

/* One passing row: the body of KernelNaive (cpu/kernel_naive.hpp:14-25).
 * Decimal64::ToValue(1,0) = 100 and Decimal64::Mul does not rescale
 * (cpu/tpch_kit.hpp:49-64), so disc_price has scale 1e-4 and charge 1e-6. */
static inline void aggregate_row(acc_t* a, int64_t qty, int64_t price, int64_t disc, int64_t tax) {
#if Q1O_MUTANT == 1
  const int64_t disc_1 = 100 - tax;
  const int64_t tax_1 = disc + 100;
#elif Q1O_MUTANT == 3
  const int64_t disc_1 = 100 + disc;
  const int64_t tax_1 = tax + 100;
#else
  const int64_t disc_1 = 100 - disc;
  const int64_t tax_1 = tax + 100;
#endif
  const int64_t disc_price = (int64_t)((uint64_t)disc_1 * (uint64_t)price); /* wrapping int64 */
#if Q1O_MUTANT == 2
  const int64_t charge = disc_price;
  (void)tax_1;
#else
  const int64_t charge = (int64_t)((uint64_t)disc_price * (uint64_t)tax_1);
#endif
  a->sum_qty = (int64_t)((uint64_t)a->sum_qty + (uint64_t)qty);
  a->sum_base_price = (int64_t)((uint64_t)a->sum_base_price + (uint64_t)price);
  a->sum_disc_price += (i128)disc_price; /* int128_add64 (cpu/common.hpp:131) */
#if Q1O_MUTANT == 8
  a->sum_charge = (i128)(int64_t)((uint64_t)a->sum_charge + (uint64_t)charge);
#else
  a->sum_charge += (i128)charge;
#endif
#if Q1O_MUTANT == 11
  a->sum_disc = (int64_t)((uint64_t)a->sum_disc + (uint64_t)disc_1);
#else
  a->sum_disc = (int64_t)((uint64_t)a->sum_disc + (uint64_t)disc);
#endif
  a->count++;
}

static inline int passes(int64_t ship, int64_t cutoff) {
#if Q1O_MUTANT == 4
  return ship < cutoff;
#else
  return ship <= cutoff;
#endif
}

/* ---- KernelNaive (cpu/kernel_naive.hpp:9-29) ------------------------------------------ */
void q1o_naive(const q1o_lineitem* li, int64_t begin, int64_t end, int32_t cutoff,
               q1o_result* out) {
  acc_t a[6];
  memset(a, 0, sizeof(a));
  int64_t unexpected = 0;
  for (int64_t i = begin; i < end; i++) {
    if (passes(li->shipdate[i], cutoff)) {
      const int s = slot_of(li->returnflag[i], li->linestatus[i]);
      if (s < 0) { unexpected++; continue; }
      aggregate_row(&a[s], li->quantity[i], li->extendedprice[i], li->discount[i], li->tax[i]);
    }
  }
  acc_to_result(a, unexpected, out);
}

/* ---- KernelHyPer<true> (cpu/kernel_hyper.hpp:73-118) ---------------------------------- */
/* Rows [begin, end) added into `a` with every operation overflow-checked; returns 1 if a
 * check fired. */
static int hyper_range(const q1o_lineitem* li, int64_t begin, int64_t end, int32_t cutoff, acc_t* a,
                       int64_t* unexpected_out) {
  int64_t unexpected = 0;
  int flag = 0;
  for (int64_t i = begin; i < end; i++) {
    if (li->shipdate[i] <= cutoff) {
      const int s = slot_of(li->returnflag[i], li->linestatus[i]);
      if (s < 0) { unexpected++; continue; }
      const int64_t disc = li->discount[i], price = li->extendedprice[i];
      long long disc_1, tax_1, disc_price, charge;
      flag |= __builtin_ssubll_overflow(100, disc, &disc_1);
      flag |= __builtin_saddll_overflow(100, li->tax[i], &tax_1);
      flag |= __builtin_smulll_overflow(disc_1, price, &disc_price);
      flag |= __builtin_smulll_overflow(disc_price, tax_1, &charge);
      long long q = a[s].sum_qty, bp = a[s].sum_base_price, d = a[s].sum_disc;
      flag |= __builtin_saddll_overflow(q, li->quantity[i], &q);
      flag |= __builtin_saddll_overflow(bp, price, &bp);
      flag |= __builtin_add_overflow(a[s].sum_disc_price, (i128)disc_price, &a[s].sum_disc_price);
      flag |= __builtin_add_overflow(a[s].sum_charge, (i128)charge, &a[s].sum_charge);
      flag |= __builtin_saddll_overflow(d, disc, &d);
      unsigned long long c = (unsigned long long)a[s].count;
      flag |= __builtin_uaddll_overflow(c, 1ull, &c);
      a[s].sum_qty = q;
      a[s].sum_base_price = bp;
      a[s].sum_disc = d;
      a[s].count = (int64_t)c;
    }
  }
  *unexpected_out += unexpected;
  return flag;
}

int q1o_hyper(const q1o_lineitem* li, int64_t begin, int64_t end, int32_t cutoff,
              q1o_result* out) {
  acc_t a[6];
  memset(a, 0, sizeof(a));
  int64_t unexpected = 0;
  const int flag = hyper_range(li, begin, end, cutoff, a, &unexpected);
  acc_to_result(a, unexpected, out);
  return flag;
}

/* Morsel-parallel KernelHyPer: workers claim morsels with fetch_add (the pattern of
 * cpu/kernel_x100.hpp:456-500), each keeps a private overflow-checked table, and the
 * tables are merged with overflow-checked adds. */
typedef struct {
  const q1o_lineitem* li;
  int64_t end;
  int32_t cutoff;
  int64_t morsel;
  _Atomic int64_t* next;
  acc_t a[6];
  int64_t unexpected;
  int flag;
} hworker_t;

static void* hyper_worker(void* arg) {
  hworker_t* w = (hworker_t*)arg;
  memset(w->a, 0, sizeof(w->a));
  w->unexpected = 0;
  w->flag = 0;
  for (;;) {
    const int64_t off = atomic_fetch_add(w->next, w->morsel);
    if (off >= w->end) break;
    const int64_t stop = off + w->morsel < w->end ? off + w->morsel : w->end;
    w->flag |= hyper_range(w->li, off, stop, w->cutoff, w->a, &w->unexpected);
  }
  return NULL;
}

int q1o_hyper_parallel(const q1o_lineitem* li, int64_t begin, int64_t end, int32_t cutoff, int threads,
                       int64_t morsel_rows, q1o_result* out) {
  if (threads < 1) threads = 1;
  if (morsel_rows < 1) morsel_rows = 10 * 1024;
  _Atomic int64_t next = begin;
  hworker_t* ws = (hworker_t*)calloc((size_t)threads, sizeof(hworker_t));
  pthread_t* ts = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
  for (int t = 0; t < threads; t++) {
    ws[t].li = li; ws[t].end = end; ws[t].cutoff = cutoff; ws[t].morsel = morsel_rows; ws[t].next = &next;
    pthread_create(&ts[t], NULL, hyper_worker, &ws[t]);
  }
  acc_t total[6];
  memset(total, 0, sizeof(total));
  int64_t unexpected = 0;
  int flag = 0;
  for (int t = 0; t < threads; t++) {
    pthread_join(ts[t], NULL);
    flag |= ws[t].flag;
    for (int g = 0; g < 6; g++) {
      long long q = total[g].sum_qty, bp = total[g].sum_base_price, d = total[g].sum_disc, c = total[g].count;
      flag |= __builtin_saddll_overflow(q, ws[t].a[g].sum_qty, &q);
      flag |= __builtin_saddll_overflow(bp, ws[t].a[g].sum_base_price, &bp);
      flag |= __builtin_saddll_overflow(d, ws[t].a[g].sum_disc, &d);
      flag |= __builtin_saddll_overflow(c, ws[t].a[g].count, &c);
      flag |= __builtin_add_overflow(total[g].sum_disc_price, ws[t].a[g].sum_disc_price, &total[g].sum_disc_price);
      flag |= __builtin_add_overflow(total[g].sum_charge, ws[t].a[g].sum_charge, &total[g].sum_charge);
      total[g].sum_qty = q;
      total[g].sum_base_price = bp;
      total[g].sum_disc = d;
      total[g].count = c;
    }
    unexpected += ws[t].unexpected;
  }
  free(ws);
  free(ts);
  acc_to_result(total, unexpected, out);
  return flag;
}

/* ---- morsel-parallel naive (pattern of cpu/kernel_x100.hpp:456-500) ------------------- */
typedef struct {
  const q1o_lineitem* li;
  int64_t end;
  int32_t cutoff;
  int64_t morsel;
  _Atomic int64_t* next;
  acc_t a[6];
  int64_t unexpected;
} worker_t;

static void* morsel_worker(void* arg) {
  worker_t* w = (worker_t*)arg;
  memset(w->a, 0, sizeof(w->a));
  w->unexpected = 0;
  for (;;) {
    const int64_t off = atomic_fetch_add(w->next, w->morsel);
    if (off >= w->end) break;
    const int64_t stop = off + w->morsel < w->end ? off + w->morsel : w->end;
    for (int64_t i = off; i < stop; i++) {
      if (passes(w->li->shipdate[i], w->cutoff)) {
        const int s = slot_of(w->li->returnflag[i], w->li->linestatus[i]);
        if (s < 0) { w->unexpected++; continue; }
        aggregate_row(&w->a[s], w->li->quantity[i], w->li->extendedprice[i], w->li->discount[i],
                      w->li->tax[i]);
      }
    }
  }
  return NULL;
}

void q1o_naive_parallel(const q1o_lineitem* li, int64_t begin, int64_t end, int32_t cutoff,
                        int threads, int64_t morsel_rows, q1o_result* out) {
  if (threads < 1) threads = 1;
  if (morsel_rows < 1) morsel_rows = 10 * 1024; /* morsel_size default, src/cpu.cpp:10 */
  _Atomic int64_t next = begin;
  worker_t* ws = (worker_t*)calloc((size_t)threads, sizeof(worker_t));
  pthread_t* ts = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
  for (int t = 0; t < threads; t++) {
    ws[t].li = li; ws[t].end = end; ws[t].cutoff = cutoff; ws[t].morsel = morsel_rows;
    ws[t].next = &next;
    pthread_create(&ts[t], NULL, morsel_worker, &ws[t]);
  }
  acc_t total[6];
  memset(total, 0, sizeof(total));
  int64_t unexpected = 0;
  for (int t = 0; t < threads; t++) {
    pthread_join(ts[t], NULL);
    for (int g = 0; g < 6; g++) { /* the workers' merge (cpu/kernel_x100.hpp:486-496) */
      total[g].sum_qty = (int64_t)((uint64_t)total[g].sum_qty + (uint64_t)ws[t].a[g].sum_qty);
      total[g].sum_base_price =
          (int64_t)((uint64_t)total[g].sum_base_price + (uint64_t)ws[t].a[g].sum_base_price);
      total[g].sum_disc = (int64_t)((uint64_t)total[g].sum_disc + (uint64_t)ws[t].a[g].sum_disc);
      total[g].sum_disc_price += ws[t].a[g].sum_disc_price;
      total[g].sum_charge += ws[t].a[g].sum_charge;
      total[g].count += ws[t].a[g].count;
    }
    unexpected += ws[t].unexpected;
  }
  free(ws);
  free(ts);
  acc_to_result(total, unexpected, out);
}

/* ---- compression (src/main.cu:773-796) and bit access (src/bit_operations.hpp) --------- */
static uint32_t encode_rf(char c) { return ((uint32_t)(c == 'R') << 1) + (uint32_t)(c == 'N'); }
static uint32_t encode_ls(char c) { return c == 'F' ? 0u : 1u; }

void q1o_compress(const q1o_lineitem* li, int64_t n, q1o_compressed* out) {
  memset(out->returnflag, 0, (size_t)((n + 15) / 16) * sizeof(uint32_t));
  memset(out->linestatus, 0, (size_t)((n + 31) / 32) * sizeof(uint32_t));
  for (int64_t i = 0; i < n; i++) {
    out->shipdate[i] = (uint16_t)(li->shipdate[i] - 727563);
    out->discount[i] = (uint8_t)li->discount[i];
    out->extendedprice[i] = (uint32_t)li->extendedprice[i];
#if Q1O_MUTANT == 9
    out->quantity[i] = (uint8_t)li->quantity[i];
#else
    out->quantity[i] = (uint8_t)(li->quantity[i] / 100);
#endif
    out->tax[i] = (uint8_t)li->tax[i];
    /* set_bit_resolution_element<1> / <0> (src/bit_operations.hpp:140-154) */
#if Q1O_MUTANT == 6
    out->returnflag[i / 16] |= encode_rf(li->returnflag[i]) << (2 * (i % 16) + 1);
#else
    out->returnflag[i / 16] |= encode_rf(li->returnflag[i]) << (2 * (i % 16));
#endif
#if Q1O_MUTANT == 7
    out->linestatus[i / 32] |= encode_ls(li->linestatus[i]) << (i % 16);
#else
    out->linestatus[i / 32] |= encode_ls(li->linestatus[i]) << (i % 32);
#endif
  }
}

uint32_t q1o_get_return_flag(const uint32_t* words, int64_t row) {
  return (words[row / 16] >> (2 * (row % 16))) & 3u;
}

uint32_t q1o_get_line_status(const uint32_t* words, int64_t row) {
  return (words[row / 32] >> (row % 32)) & 1u;
}

/* Accumulate a compressed row the way the GPU compressed kernels do
 * (src/kernels/ht_in_registers_per_thread.cuh:107-133): slot = rf << 1 | ls of the packed
 * codes, quantity in whole units. */
static inline void aggregate_compressed(acc_t* a, const q1o_compressed* c, int64_t i) {
  const uint32_t rf = q1o_get_return_flag(c->returnflag, i);
  const uint32_t ls = q1o_get_line_status(c->linestatus, i);
#if Q1O_MUTANT == 5
  const uint32_t s = ls * 3 + rf;
#else
  const uint32_t s = rf << 1 | ls;
#endif
  if (s < 6)
    aggregate_row(&a[s], c->quantity[i], c->extendedprice[i], c->discount[i], c->tax[i]);
}

void q1o_compressed_kernel(const q1o_compressed* c, int64_t begin, int64_t end, int32_t cutoff,
                           q1o_result* out) {
  acc_t a[6];
  memset(a, 0, sizeof(a));
  const int64_t ccut = (int64_t)cutoff - 727563; /* compressed_threshold_ship_date */
  for (int64_t i = begin; i < end; i++)
    if (passes(c->shipdate[i], ccut)) aggregate_compressed(a, c, i);
  acc_to_result(a, 0, out);
}

/* The compressed-kernel semantics morsel-parallel, in the pattern of Morsel::Query
 * (cpu/kernel_x100.hpp:456-500): workers claim fixed-size morsels of [begin, end) with
 * fetch_add, aggregate them into private tables, and the tables are added up exactly.
 * Used to check large device runs (hundreds of millions of rows) in seconds. */
typedef struct {
  const q1o_compressed* c;
  int64_t end, ccut, morsel;
  _Atomic int64_t* next;
  acc_t a[6];
} cworker_t;

static void* compressed_worker(void* arg) {
  cworker_t* w = (cworker_t*)arg;
  memset(w->a, 0, sizeof(w->a));
  for (;;) {
    const int64_t off = atomic_fetch_add(w->next, w->morsel);
    if (off >= w->end) break;
    const int64_t stop = off + w->morsel < w->end ? off + w->morsel : w->end;
#if Q1O_MUTANT == 14
    if (stop - off < w->morsel) break; /* planted: the last partial morsel is dropped */
#endif
    for (int64_t i = off; i < stop; i++)
      if (passes(w->c->shipdate[i], w->ccut)) aggregate_compressed(w->a, w->c, i);
  }
  return NULL;
}

void q1o_compressed_parallel(const q1o_compressed* c, int64_t begin, int64_t end, int32_t cutoff,
                             int threads, int64_t morsel_rows, q1o_result* out) {
  if (threads < 1) threads = 1;
  if (morsel_rows < 1) morsel_rows = 1 << 20;
  _Atomic int64_t next = begin;
  cworker_t* ws = (cworker_t*)calloc((size_t)threads, sizeof(cworker_t));
  pthread_t* ts = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
  for (int t = 0; t < threads; t++) {
    ws[t].c = c; ws[t].end = end; ws[t].ccut = (int64_t)cutoff - 727563; ws[t].morsel = morsel_rows;
    ws[t].next = &next;
    pthread_create(&ts[t], NULL, compressed_worker, &ws[t]);
  }
  acc_t total[6];
  memset(total, 0, sizeof(total));
  for (int t = 0; t < threads; t++) {
    pthread_join(ts[t], NULL);
    for (int g = 0; g < 6; g++) {
      total[g].sum_qty += ws[t].a[g].sum_qty;
      total[g].sum_base_price += ws[t].a[g].sum_base_price;
      total[g].sum_disc += ws[t].a[g].sum_disc;
      total[g].sum_disc_price += ws[t].a[g].sum_disc_price;
      total[g].sum_charge += ws[t].a[g].sum_charge;
      total[g].count += ws[t].a[g].count;
    }
  }
  free(ws);
  free(ts);
  acc_to_result(total, 0, out);
}

/* ---- filter pushdown producer (src/main.cu:362-389) and consumer ----------------------- */
void q1o_precompute_filter(const uint16_t* shipdate, int64_t n, int32_t cutoff, uint32_t* out) {
  const int64_t ccut = (int64_t)cutoff - 727563;
  const int64_t full = n - n % 32;
  for (int64_t i = 0; i < full; i += 32) {
    uint32_t w = 0;
    for (int j = 0; j < 32; j++) w |= (uint32_t)(shipdate[i + j] <= ccut) << j;
    out[i / 32] = w;
  }
#if Q1O_MUTANT != 10
  if (n > full) {
    uint32_t w = 0;
    for (int64_t j = 0; full + j < n; j++) w |= (uint32_t)(shipdate[full + j] <= ccut) << j;
    out[n / 32] = w;
  }
#endif
}

void q1o_pushdown_kernel(const uint32_t* filter, const q1o_compressed* c, int64_t begin,
                         int64_t end, q1o_result* out) {
  acc_t a[6];
  memset(a, 0, sizeof(a));
  for (int64_t i = begin; i < end; i++)
    if ((filter[i / 32] >> (i % 32)) & 1u) aggregate_compressed(a, c, i); /* get_bit */
  acc_to_result(a, 0, out);
}

/* ---- ComprData (cpu/common.cpp:113-137, the -DGPU variant) ---------------------------- */
void q1o_compact_init(const q1o_lineitem* li, int64_t n, q1o_compact* out) {
  for (int64_t i = 0; i < n; i++) {
    out->shipdate[i] = (int16_t)(li->shipdate[i] - 727563);
    out->returnflag[i] = (int8_t)li->returnflag[i];
    out->linestatus[i] = (int8_t)li->linestatus[i];
    out->discount[i] = (int8_t)li->discount[i];
    out->tax[i] = (int8_t)li->tax[i];
    out->extendedprice[i] = (int32_t)li->extendedprice[i];
    out->quantity[i] = (int16_t)li->quantity[i];
  }
}

/* ---- KernelX100::task (cpu/kernel_x100.hpp:140-402) with the k1Step aggregation ------- */
/* One vector of at most 1024 rows [off, off + n) of the compact columns.  The primitives'
 * types are kept: select on int16 dates (vectorized.cpp:504-506), disc_1 and tax_1 in int8
 * (:544-554), disc_price in int32 (:557-564), charge in int64 (:447-498), and sum_disc
 * accumulates disc_1 (kernel_x100.hpp:320). */
enum { kX100Vector = 1024 };

static void x100_vector(const q1o_compact* c, int64_t off, int64_t n, int16_t date, acc_t* a,
                        int64_t* unexpected) {
  uint16_t sel[kX100Vector];
  int num = 0;
  for (int64_t i = 0; i < n; i++) /* select_int16_t: a[i] <= b */
    if (c->shipdate[off + i] <= date) sel[num++] = (uint16_t)i;
  const int8_t one = 100;                            /* Decimal64::ToValue(1, 0) */
  for (int k = 0; k < num; k++) {
    const int64_t i = off + sel[k];
    const int s = slot_of((char)c->returnflag[i], (char)c->linestatus[i]);
    if (s < 0) { (*unexpected)++; continue; }
    const int8_t disc_1 = (int8_t)(one - c->discount[i]);                       /* map_disc_1 */
    const int8_t tax_1 = (int8_t)(one + c->tax[i]);                             /* map_tax_1 */
    const int32_t disc_price = (int32_t)((uint32_t)(int32_t)disc_1 * (uint32_t)c->extendedprice[i]);
    const int64_t charge = (int64_t)disc_price * tax_1;                         /* map_charge */
    acc_t* g = &a[s];
    g->sum_qty = (int64_t)((uint64_t)g->sum_qty + (uint64_t)(int64_t)c->quantity[i]);
    g->sum_base_price = (int64_t)((uint64_t)g->sum_base_price + (uint64_t)(int64_t)c->extendedprice[i]);
    g->sum_disc_price += (i128)disc_price;
    g->sum_charge += (i128)charge;
    g->sum_disc = (int64_t)((uint64_t)g->sum_disc + (uint64_t)(int64_t)disc_1);
    g->count++;
  }
}

typedef struct {
  const q1o_compact* c;
  int64_t end;
  int16_t date;
  int64_t morsel;
  _Atomic int64_t* next;
  acc_t a[6];
  int64_t unexpected;
} x100_worker_t;

static void* x100_worker(void* arg) {
  x100_worker_t* w = (x100_worker_t*)arg;
  memset(w->a, 0, sizeof(w->a));
  w->unexpected = 0;
  for (;;) {
    const int64_t off = atomic_fetch_add(w->next, w->morsel);
    if (off >= w->end) break;
    const int64_t stop = off + w->morsel < w->end ? off + w->morsel : w->end;
    for (int64_t v = off; v < stop; v += kX100Vector) {
      const int64_t n = stop - v < kX100Vector ? stop - v : kX100Vector;
#if Q1O_MUTANT == 12
      if (n < kX100Vector) break; /* planted: the last partial vector is dropped */
#endif
      x100_vector(w->c, v, n, w->date, w->a, &w->unexpected);
    }
  }
  return NULL;
}

void q1o_x100(const q1o_compact* c, int64_t begin, int64_t end, int32_t cutoff, int threads,
              int64_t morsel_rows, q1o_result* out) {
  if (threads < 1) threads = 1;
  if (morsel_rows < 1) morsel_rows = 10 * 1024;
  /* `const int16_t date = cmp.dte_val - 727563` (kernel_x100.hpp:147-148) */
  const int16_t date = (int16_t)(cutoff - 727563);
  _Atomic int64_t next = begin;
  x100_worker_t* ws = (x100_worker_t*)calloc((size_t)threads, sizeof(x100_worker_t));
  pthread_t* ts = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
  for (int t = 0; t < threads; t++) {
    ws[t].c = c; ws[t].end = end; ws[t].date = date; ws[t].morsel = morsel_rows; ws[t].next = &next;
    pthread_create(&ts[t], NULL, x100_worker, &ws[t]);
  }
  acc_t total[6];
  memset(total, 0, sizeof(total));
  int64_t unexpected = 0;
  for (int t = 0; t < threads; t++) {
    pthread_join(ts[t], NULL);
    for (int g = 0; g < 6; g++) {
      total[g].sum_qty = (int64_t)((uint64_t)total[g].sum_qty + (uint64_t)ws[t].a[g].sum_qty);
      total[g].sum_base_price =
          (int64_t)((uint64_t)total[g].sum_base_price + (uint64_t)ws[t].a[g].sum_base_price);
      total[g].sum_disc = (int64_t)((uint64_t)total[g].sum_disc + (uint64_t)ws[t].a[g].sum_disc);
      total[g].sum_disc_price += ws[t].a[g].sum_disc_price;
      total[g].sum_charge += ws[t].a[g].sum_charge;
      total[g].count += ws[t].a[g].count;
    }
    unexpected += ws[t].unexpected;
  }
  free(ws);
  free(ts);
  acc_to_result(total, unexpected, out);
}

/* ---- AVX512<false, false, true> (cpu/kernel_avx512.hpp:11-194), lane by lane ----------- */
/* Steps of 16 rows in two halves of 8; row j of a half updates the 8-way replicated table
 * entry gid * 8 + j (gid = rf << 8 | ls, make_index :113-119).  The 128-bit sums are
 * (lo, hi) word pairs updated as add64_to_int128 does (cpu/common.hpp:102-117): lo += v,
 * then hi += 1 when the new lo is below v in a *signed* comparison.  The final per-group
 * result adds the 8 entries (cpu/q1.cpp:86-101): 64-bit sums plainly, 128-bit sums as
 * (hi << 64) | (int128)lo with lo sign-extended.  Below 2^63 per entry this equals the
 * naive sums exactly; above, the reference's carry and combine rules are kept as they are. */
enum { kAvxLanes = 8, kAvxGroups = 1 << 16 };

typedef struct {
  int64_t count, sum_qty, sum_base_price, sum_disc;
  int64_t dp_lo, dp_hi, ch_lo, ch_hi;
} avx_entry_t;

static inline void avx_add128(int64_t* hi, int64_t* lo, int64_t v) {
  const int64_t nlo = (int64_t)((uint64_t)*lo + (uint64_t)v);
  if (nlo < v) *hi = (int64_t)((uint64_t)*hi + 1u);  /* _mm512_cmplt_epi64_mask: signed */
  *lo = nlo;
}

void q1o_avx512(const q1o_lineitem* li, int64_t begin, int64_t end, int32_t cutoff, q1o_result* out) {
  avx_entry_t* t = (avx_entry_t*)calloc((size_t)kAvxGroups * kAvxLanes, sizeof(avx_entry_t));
  for (int64_t off = begin; off < end; off += 16) {
    for (int j = 0; j < 16; j++) {
      const int64_t i = off + j;
      if (i >= end || li->shipdate[i] > cutoff) continue; /* the select and cardinality masks */
      const int64_t disc = li->discount[i], price = li->extendedprice[i];
      const int64_t disc_1 = 100 - disc, tax_1 = 100 + li->tax[i];
      const int64_t disc_price = (int64_t)((uint64_t)disc_1 * (uint64_t)price);  /* mul_64_64_64 */
      const int64_t charge = (int64_t)((uint64_t)disc_price * (uint64_t)tax_1);
      const uint32_t gid = (uint32_t)(uint8_t)li->returnflag[i] << 8 | (uint8_t)li->linestatus[i];
      avx_entry_t* e = &t[(size_t)gid * kAvxLanes + (j & 7)];
      e->count++;
      e->sum_qty = (int64_t)((uint64_t)e->sum_qty + (uint64_t)li->quantity[i]);
      e->sum_base_price = (int64_t)((uint64_t)e->sum_base_price + (uint64_t)price);
      e->sum_disc = (int64_t)((uint64_t)e->sum_disc + (uint64_t)disc);
      avx_add128(&e->dp_hi, &e->dp_lo, disc_price);
      avx_add128(&e->ch_hi, &e->ch_lo, charge);
    }
  }
  memset(out, 0, sizeof(*out));
  for (uint32_t gid = 0; gid < kAvxGroups; gid++) {
    acc_t a;
    memset(&a, 0, sizeof(a));
#if Q1O_MUTANT == 13
    const int lanes = kAvxLanes - 1; /* planted: the last lane is left out of the final sum */
#else
    const int lanes = kAvxLanes;
#endif
    for (int l = 0; l < lanes; l++) {
      const avx_entry_t* e = &t[(size_t)gid * kAvxLanes + l];
      a.count += e->count;
      a.sum_qty = (int64_t)((uint64_t)a.sum_qty + (uint64_t)e->sum_qty);
      a.sum_base_price = (int64_t)((uint64_t)a.sum_base_price + (uint64_t)e->sum_base_price);
      a.sum_disc = (int64_t)((uint64_t)a.sum_disc + (uint64_t)e->sum_disc);
      a.sum_disc_price += (i128)((unsigned __int128)(uint64_t)e->dp_hi << 64) | (i128)e->dp_lo;
      a.sum_charge += (i128)((unsigned __int128)(uint64_t)e->ch_hi << 64) | (i128)e->ch_lo;
    }
    if (a.count <= 0) continue;
    const int s = slot_of((char)(gid >> 8), (char)(gid & 0xFF));
    if (s < 0) { out->unexpected_keys += a.count; continue; }
    q1o_group* o = &out->g[s];
    o->sum_qty = a.sum_qty;
    o->sum_base_price = a.sum_base_price;
    o->sum_disc_price_lo = (uint64_t)a.sum_disc_price;
    o->sum_disc_price_hi = (int64_t)(a.sum_disc_price >> 64);
    o->sum_charge_lo = (uint64_t)a.sum_charge;
    o->sum_charge_hi = (int64_t)(a.sum_charge >> 64);
    o->sum_disc = a.sum_disc;
    o->count = a.count;
  }
  free(t);
}
