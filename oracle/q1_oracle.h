/*
 * q1_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's Q1 algorithm, the test reference for the HIP path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product never links, imports or calls anything here.
 *
 * The reference's own CPU kernels (cpu/kernel_naive.hpp, cpu/kernel_hyper.hpp) cannot be
 * compiled in this environment without writing a stand-in for Boost.Spirit
 * (Date::parse in cpu/tpch_kit.cpp:124-139 is reached from BaseKernel's constructor,
 * cpu/common.cpp:59-60), so every function here is pinned instead by closed forms,
 * a library date routine, invariants, brute force on tiny inputs and planted-mistake
 * builds (tests/test_oracle.py).
 */
#ifndef Q1_ORACLE_H_
#define Q1_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* AggrHashTable (cpu/common.hpp:158-165), with the int128 sums split into words so that
 * ctypes can read them. */
typedef struct {
  int64_t sum_qty;
  int64_t sum_base_price;
  uint64_t sum_disc_price_lo;
  int64_t sum_disc_price_hi;
  uint64_t sum_charge_lo;
  int64_t sum_charge_hi;
  int64_t sum_disc;
  int64_t count;
} q1o_group;

/* Result keyed like the GPU kernels: slot = encode_rf << 1 | encode_ls
 * (src/data_types.h:52-80).  `unexpected_keys` counts passing rows whose
 * (returnflag, linestatus) pair is not one of A/N/R x F/O. */
typedef struct {
  q1o_group g[6];
  int64_t unexpected_keys;
} q1o_result;

/* The reference lineitem columns (cpu/tpch_kit.hpp:135-142). */
typedef struct {
  const int32_t* shipdate;
  const int64_t* quantity;      /* Decimal(15,2) x100 */
  const int64_t* extendedprice; /* x100 */
  const int64_t* discount;      /* x100 */
  const int64_t* tax;           /* x100 */
  const char* returnflag;
  const char* linestatus;
} q1o_lineitem;

/* Compressed columns (src/main.cu:752-761; compressed:: types in src/data_types.h:35-44). */
typedef struct {
  uint16_t* shipdate;       /* shipdate - 727563 */
  uint8_t* quantity;        /* quantity / 100 */
  uint32_t* extendedprice;
  uint8_t* discount;
  uint8_t* tax;
  uint32_t* returnflag;     /* 2 bits per row, 16 rows per word, LSB first */
  uint32_t* linestatus;     /* 1 bit per row, 32 rows per word, LSB first */
} q1o_compressed;

/* todate (cpu/monetdb.cpp:61-77). */
int32_t q1o_todate(int day, int month, int year);

/* KernelNaive::operator() (cpu/kernel_naive.hpp:9-29) over rows [begin, end) with the
 * filter `shipdate <= cutoff` (the reference fixes cutoff = todate(1998-12-01) - 90,
 * cpu/common.cpp:59-60). */
void q1o_naive(const q1o_lineitem* li, int64_t begin, int64_t end, int32_t cutoff,
               q1o_result* out);

/* KernelHyPer<true>::operator() (cpu/kernel_hyper.hpp:73-118): the naive semantics with
 * every operation overflow-checked; returns 1 if any check fired (where the reference
 * calls handle_overflow(), cpu/common.cpp:7-11), else 0. */
int q1o_hyper(const q1o_lineitem* li, int64_t begin, int64_t end, int32_t cutoff,
              q1o_result* out);

/* Morsel-parallel KernelHyPer<true> on `threads` threads: private overflow-checked tables
 * per worker, merged with overflow-checked adds; returns 1 if any check fired. */
int q1o_hyper_parallel(const q1o_lineitem* li, int64_t begin, int64_t end, int32_t cutoff, int threads,
                       int64_t morsel_rows, q1o_result* out);

/* Morsel-parallel KernelNaive on `threads` threads (pattern of Morsel::Query,
 * cpu/kernel_x100.hpp:456-500: workers claim fixed-size morsels with fetch_add and
 * merge their private tables).  Used as the timed CPU baseline. */
void q1o_naive_parallel(const q1o_lineitem* li, int64_t begin, int64_t end, int32_t cutoff,
                        int threads, int64_t morsel_rows, q1o_result* out);

/* The compression loop of main (src/main.cu:773-796) for rows [0, n).  The packed
 * word arrays must hold ceil(n/16) resp. ceil(n/32) words; they are zeroed first. */
void q1o_compress(const q1o_lineitem* li, int64_t n, q1o_compressed* out);

/* Bit decode get_bit_resolution_element (src/bit_operations.hpp:126-138). */
uint32_t q1o_get_return_flag(const uint32_t* words, int64_t row);
uint32_t q1o_get_line_status(const uint32_t* words, int64_t row);

/* The compressed-kernel semantics (src/kernels/ht_in_registers_per_thread.cuh:107-133
 * and siblings) over rows [begin, end): filter `shipdate <= cutoff - 727563`,
 * quantity summed in whole units as the GPU kernels do. */
void q1o_compressed_kernel(const q1o_compressed* c, int64_t begin, int64_t end, int32_t cutoff,
                           q1o_result* out);

/* q1o_compressed_kernel morsel-parallel on `threads` threads (the morsel pattern of
 * cpu/kernel_x100.hpp:456-500), for checking device runs over hundreds of millions of
 * rows. */
void q1o_compressed_parallel(const q1o_compressed* c, int64_t begin, int64_t end, int32_t cutoff,
                             int threads, int64_t morsel_rows, q1o_result* out);

/* precompute_filter_for_table_chunk (src/main.cu:362-389) with a parameter cutoff, over
 * rows [0, n); writes ceil(n/32) words. */
void q1o_precompute_filter(const uint16_t* shipdate, int64_t n, int32_t cutoff, uint32_t* out);

/* The filter-pushdown kernel semantics (..._filter_pushdown_compressed, e.g.
 * src/kernels/ht_in_registers_per_thread.cuh:176-203) over rows [begin, end). */
void q1o_pushdown_kernel(const uint32_t* filter, const q1o_compressed* c, int64_t begin,
                         int64_t end, q1o_result* out);

/* ComprData (cpu/common.cpp:113-137, the -DGPU variant): the compact columns the X100
 * kernel reads; quantity and prices keep the x100 scale. */
typedef struct {
  int16_t* shipdate;        /* shipdate - 727563 */
  int8_t* returnflag;
  int8_t* linestatus;
  int8_t* discount;
  int8_t* tax;
  int32_t* extendedprice;
  int16_t* quantity;
} q1o_compact;
void q1o_compact_init(const q1o_lineitem* li, int64_t n, q1o_compact* out);

/* KernelX100::task (cpu/kernel_x100.hpp:140-402) with the k1Step aggregation, morsel-
 * parallel as Morsel::Query (:456-500): vectors of 1024 rows, select on int16 dates,
 * int8 disc_1 / tax_1, int32 disc_price, int64 charge, and sum_disc = sum of disc_1
 * (the reference's quirk, kernel_x100.hpp:320). */
void q1o_x100(const q1o_compact* c, int64_t begin, int64_t end, int32_t cutoff, int threads,
              int64_t morsel_rows, q1o_result* out);

/* AVX512<false, false, true> (cpu/kernel_avx512.hpp:11-194) lane by lane: an 8-way
 * replicated table per (returnflag << 8 | linestatus), the reference's 128-bit carry rule
 * (cpu/common.hpp:102-117) and final 8-lane sum (cpu/q1.cpp:86-101). */
void q1o_avx512(const q1o_lineitem* li, int64_t begin, int64_t end, int32_t cutoff, q1o_result* out);

#ifdef __cplusplus
}
#endif

#endif /* Q1_ORACLE_H_ */
