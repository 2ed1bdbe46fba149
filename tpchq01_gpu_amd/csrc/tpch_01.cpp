// tpch_01: the TPC-H Query 1 host driver (the reference's src/main.cu:738-916), calling the
// gfx950 path only through the C ABI in include/q1_hip.h.
//
// Options kept from the reference (src/main.cu:255-360; README spellings accepted too):
//   --apply-compression --use-filter-pushdown --use-coprocessing --print-results --device
//   --hash-table-placement=in_registers|in_registers_per_thread|local_mem|
//                          shared_mem_per_thread|global   (or in-registers, local-mem,
//                          per-thread-shared-mem; plus auto)
//   --sf=N / --scale-factor=N   --runs=N   --cpu-fraction=F   --streams=N
//   --threads-per-block=N   --tuples-per-thread=N   --tuples-per-kernel-launch=N
//   (--tuples-per-kernel=N)
// Added:
//   --rows=N             row count (default: sf x 6001215, the SF 1 lineitem cardinality)
//   --device-resident    load the columns into HBM once and time only the query
//   --data-dir=DIR       read DIR/<sf>/{shipdate,...}.bin (the reference's column cache) or
//                        parse DIR/<sf>/lineitem.tbl and write the cache; otherwise the
//                        seeded synthetic generator supplies the rows
//   --seed=N --cutoff=YYYY-MM-DD --cpu-threads=N
//   --filter-once        with --use-filter-pushdown: build the filter bitmap once, before
//                        the timed runs (off the critical path), instead of in every run
//   --ingest-only        with --data-dir: parse DIR/<sf>/lineitem.tbl, (re)write the
//                        column cache and exit, without touching the GPU
// Each run's wall time goes to results.csv (src/main.cu:881-905).
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/q1_hip.h"

namespace {

struct Params {
  double scale_factor = 1.0;
  std::string placement = "in_registers_per_thread";  // the reference default (constants.hpp:141)
  bool print_results = false;
  bool filter_pushdown = false;
  bool compression = false;
  bool coprocessing = false;
  bool device_resident = false;
  int streams = 4;
  uint32_t threads_per_block = 0;
  uint32_t tuples_per_thread = 0;
  uint64_t tuples_per_launch = 1u << 21;
  int runs = 5;
  double cpu_fraction = 0.5;
  uint64_t rows = 0;
  std::string data_dir;
  uint64_t seed = 0x51C00000ull;
  int32_t cutoff = 729999;
  int cpu_threads = 0;
  bool user_threads = false;
  bool filter_once = false;
  bool ingest_only = false;
};

[[noreturn]] void usage(const char* argv0) {
  std::fprintf(stderr,
               "Usage: %s [args]\n"
               "   --apply-compression\n   --print-results\n   --use-filter-pushdown\n   --use-coprocessing\n"
               "   --device-resident\n   --device\n"
               "   --hash-table-placement=[default:in_registers_per_thread]\n"
               "     (one of: in_registers, in_registers_per_thread, local_mem, shared_mem_per_thread, global, auto;\n"
               "      also in-registers, local-mem, per-thread-shared-mem)\n"
               "   --runs=[default:5]\n   --sf=|--scale-factor=[default:1]\n   --rows=N\n"
               "   --cpu-fraction=[default:0.5]\n   --streams=[default:4]\n"
               "   --threads-per-block=N\n   --tuples-per-thread=N\n   --tuples-per-kernel-launch=[default:2097152]\n"
               "   --data-dir=DIR\n   --seed=N\n   --cutoff=YYYY-MM-DD\n   --cpu-threads=N\n   --filter-once\n"
               "   --ingest-only\n",
               argv0);
  std::exit(EXIT_FAILURE);
}

int placement_code(const std::string& s) {
  if (s == "in_registers" || s == "in-registers") return Q1_PLACE_IN_REGISTERS;
  if (s == "in_registers_per_thread" || s == "in-registers-per-thread") return Q1_PLACE_IN_REGISTERS_PER_THREAD;
  if (s == "local_mem" || s == "local-mem") return Q1_PLACE_LOCAL_MEM;
  if (s == "shared_mem_per_thread" || s == "per-thread-shared-mem") return Q1_PLACE_PER_THREAD_SHARED_MEM;
  if (s == "global") return Q1_PLACE_GLOBAL;
  if (s == "auto") return Q1_PLACE_AUTO;
  return -1;
}

// "YYYY-MM-DD" exactly (the .tbl date form), a real calendar day -> MonetDB day number.
bool parse_date(const std::string& s, int32_t* out) {
  if (s.size() != 10 || s[4] != '-' || s[7] != '-') return false;
  for (int i : {0, 1, 2, 3, 5, 6, 8, 9})
    if (s[i] < '0' || s[i] > '9') return false;
  const int y = std::stoi(s.substr(0, 4)), m = std::stoi(s.substr(5, 2)), d = std::stoi(s.substr(8, 2));
  static const int kDays[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
  const bool leap = (y % 4 == 0 && y % 100 != 0) || y % 400 == 0;
  if (m < 1 || m > 12 || d < 1 || d > kDays[m - 1] + (m == 2 && leap)) return false;
  *out = q1_day_number(y, m, d);
  return true;
}

void print_device(q1_ctx* ctx) {
  q1_device_info d;
  if (q1_device_info_get(ctx, &d) != Q1_OK) return;
  std::printf("Device name: %s\nArchitecture: %s\nNumber of compute units: %d\nClock: %d kHz\n"
              "Memory clock: %d kHz\nMemory bus width: %d bits\nGlobal memory: %llu bytes\n"
              "LDS per block: %llu bytes\nWavefront size: %d\nMax threads per block: %d\n",
              d.name, d.arch, d.compute_units, d.clock_khz, d.memory_clock_khz, d.memory_bus_width,
              (unsigned long long)d.total_memory, (unsigned long long)d.lds_per_block, d.wavefront_size,
              d.max_threads_per_block);
}

Params parse(int argc, char** argv, bool* want_device) {
  Params p;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    if (a.rfind("--", 0) != 0) usage(argv[0]);
    a = a.substr(2);
    const size_t eq = a.find('=');
    const std::string name = a.substr(0, eq), value = eq == std::string::npos ? "" : a.substr(eq + 1);
    try {
      if (name == "device") *want_device = true;
      else if (name == "use-coprocessing") p.coprocessing = true;
      else if (name == "apply-compression") p.compression = true;
      else if (name == "use-filter-pushdown") p.filter_pushdown = true;
      else if (name == "print-results") p.print_results = true;
      else if (name == "device-resident") p.device_resident = true;
      else if (name == "filter-once") p.filter_once = true;
      else if (name == "ingest-only") p.ingest_only = true;
      else if (name == "scale-factor" || name == "sf") {
        p.scale_factor = std::stod(value);
        if (p.scale_factor < 0.001) throw std::invalid_argument("Invalid scale factor " + value);
      } else if (name == "hash-table-placement") {
        if (placement_code(value) < 0) throw std::invalid_argument("No kernel variant named \"" + value + "\" is available");
        p.placement = value;
      } else if (name == "streams") {
        p.streams = std::stoi(value);
      } else if (name == "tuples-per-thread" || name == "tuples_per_thread") {
        p.tuples_per_thread = (uint32_t)std::stoul(value);
      } else if (name == "threads-per-block") {
        p.threads_per_block = (uint32_t)std::stoul(value);
        p.user_threads = true;
        if (p.threads_per_block != 64 && p.threads_per_block != 128 && p.threads_per_block != 256)
          throw std::invalid_argument("threads per block must be 64, 128 or 256 (whole wavefronts of 64)");
      } else if (name == "tuples-per-kernel-launch" || name == "tuples-per-kernel") {
        p.tuples_per_launch = std::stoull(value);
      } else if (name == "runs") {
        p.runs = std::stoi(value);
        if (p.runs <= 0) throw std::invalid_argument("Number of runs must be positive");
      } else if (name == "cpu-fraction") {
        p.cpu_fraction = std::stod(value);
        if (p.cpu_fraction < 0 || p.cpu_fraction > 1.0)
          throw std::invalid_argument("The fraction of aggregation work be performed by the CPU must be in the range 0.0 - 1.0");
      } else if (name == "rows") {
        p.rows = std::stoull(value);
      } else if (name == "data-dir") {
        p.data_dir = value;
      } else if (name == "seed") {
        p.seed = std::stoull(value, nullptr, 0);
      } else if (name == "cutoff") {
        if (!parse_date(value, &p.cutoff)) throw std::invalid_argument("Cannot parse the date passed after --cutoff=");
      } else if (name == "cpu-threads") {
        p.cpu_threads = std::stoi(value);
      } else {
        usage(argv[0]);
      }
    } catch (const std::invalid_argument& e) {
      std::fprintf(stderr, "%s\n", e.what());
      std::exit(EXIT_FAILURE);
    } catch (const std::out_of_range&) {
      std::fprintf(stderr, "Value out of range for --%s\n", name.c_str());
      std::exit(EXIT_FAILURE);
    }
  }
  if (p.filter_pushdown && !p.compression) {
    std::fprintf(stderr,
                 "Filter precomputation is only currently supported when compression is applied; "
                 "invoke with \"--apply-compression\".\n");
    std::exit(EXIT_FAILURE);
  }
  if (p.streams < 1 || p.streams > 16) {
    std::fprintf(stderr, "--streams must be in 1..16\n");
    std::exit(EXIT_FAILURE);
  }
  if (p.tuples_per_launch == 0 || p.tuples_per_launch % 32) {
    std::fprintf(stderr, "--tuples-per-kernel-launch must be a positive multiple of 32\n");
    std::exit(EXIT_FAILURE);
  }
  if (p.ingest_only && p.data_dir.empty()) {
    std::fprintf(stderr, "--ingest-only needs --data-dir\n");
    std::exit(EXIT_FAILURE);
  }
  if (p.rows == 0) p.rows = (uint64_t)(6001215.0 * p.scale_factor);
  return p;
}

// ---- host columns --------------------------------------------------------------------
// Host columns of one layout: pinned through the context, or plain pageable memory when
// there is no context (--ingest-only, which never touches the GPU).
struct HostColumns {
  q1_layout layout;
  uint64_t rows = 0;
  std::vector<void*> bufs;
  q1_columns cols{};
  q1_ctx* ctx = nullptr;
  HostColumns(q1_ctx* c, q1_layout l, uint64_t n) : layout(l), rows(n), ctx(c) {
    uint64_t b[8];
    q1_layout_column_bytes(l, n, b);
    const void** slots[8] = {&cols.ship_date, &cols.quantity, &cols.extended_price, &cols.discount,
                             &cols.tax, &cols.return_flag, &cols.line_status, &cols.filter};
    for (int i = 0; i < 8; i++) {
      if (!b[i]) continue;
      void* p = ctx ? q1_host_alloc(ctx, b[i]) : std::aligned_alloc(64, (b[i] + 63) / 64 * 64);
      if (!p) throw std::runtime_error("host column allocation failed");
      bufs.push_back(p);
      *slots[i] = p;
    }
  }
  ~HostColumns() {
    for (void* p : bufs) {
      if (ctx)
        q1_host_free(ctx, p);
      else
        std::free(p);
    }
  }
};

bool file_exists(const std::string& f) {
  struct stat st;
  return stat(f.c_str(), &st) == 0;
}

std::string sf_dir(const Params& p) { return p.data_dir + "/" + std::to_string(p.scale_factor); }

const char* kBinNames[7] = {"shipdate.bin", "quantity.bin", "extendedprice.bin", "discount.bin",
                            "tax.bin", "returnflag.bin", "linestatus.bin"};

// The reference's raw column cache (src/main.cu:405-441, file_access.hpp:19-61): one file
// per column in the reference types.
std::unique_ptr<HostColumns> load_bin_cache(q1_ctx* ctx, const Params& p) {
  const std::string dir = sf_dir(p);
  struct stat st;
  if (stat((dir + "/shipdate.bin").c_str(), &st) != 0) return nullptr;
  const uint64_t rows = (uint64_t)st.st_size / 4;
  if (rows == 0) throw std::runtime_error("The lineitem table column cardinality should not be 0");
  auto h = std::make_unique<HostColumns>(ctx, Q1_LAYOUT_REFERENCE, rows);
  uint64_t b[8];
  q1_layout_column_bytes(Q1_LAYOUT_REFERENCE, rows, b);
  const void* dst[7] = {h->cols.ship_date, h->cols.quantity, h->cols.extended_price, h->cols.discount,
                        h->cols.tax, h->cols.return_flag, h->cols.line_status};
  for (int i = 0; i < 7; i++) {
    std::ifstream f(dir + "/" + kBinNames[i], std::ios::binary);
    if (!f || !f.read((char*)dst[i], (std::streamsize)b[i]))
      throw std::runtime_error("cannot read column file " + dir + "/" + kBinNames[i]);
  }
  return h;
}

void write_bin_cache(const Params& p, const HostColumns& h) {
  const std::string dir = sf_dir(p);
  uint64_t b[8];
  q1_layout_column_bytes(Q1_LAYOUT_REFERENCE, h.rows, b);
  const void* src[7] = {h.cols.ship_date, h.cols.quantity, h.cols.extended_price, h.cols.discount,
                        h.cols.tax, h.cols.return_flag, h.cols.line_status};
  for (int i = 0; i < 7; i++) {
    std::ofstream f(dir + "/" + kBinNames[i], std::ios::binary);
    f.write((const char*)src[i], (std::streamsize)b[i]);
    if (!f) throw std::runtime_error("cannot write column file " + dir + "/" + kBinNames[i]);
  }
}

// Decimal(15,2) text -> value x100 (cpu/tpch_kit.hpp:41-82 semantics for "int.frac"):
// an optional '-', at least one digit, then optionally '.' and one or two digits
// ("17" -> 1700, "0.1" -> 10, "21168.23" -> 2116823).  False for anything else.
bool parse_decimal(const char* s, const char* e, int64_t* out) {
  bool neg = false;
  if (s < e && *s == '-') { neg = true; s++; }
  int64_t ip = 0, fp = 0;
  int id = 0, fd = 0;
  for (; s < e && *s >= '0' && *s <= '9'; s++, id++) {
    if (id >= 15) return false;
    ip = ip * 10 + (*s - '0');
  }
  if (id == 0) return false;
  if (s < e && *s == '.') {
    s++;
    for (; s < e && *s >= '0' && *s <= '9'; s++, fd++) {
      if (fd >= 2) return false;
      fp = fp * 10 + (*s - '0');
    }
    if (fd == 0) return false;
  }
  if (s != e) return false;
  if (fd == 1) fp *= 10;
  const int64_t v = ip * 100 + fp;
  *out = neg ? -v : v;
  return true;
}

// lineitem.tbl (dbgen, '|'-separated): fields 5-11 are quantity, extendedprice, discount,
// tax, returnflag, linestatus, shipdate (cpu/tpch_kit.cpp:156-176).
std::unique_ptr<HostColumns> parse_tbl(q1_ctx* ctx, const std::string& path) {
  std::ifstream in(path);
  if (!in) return nullptr;
  std::vector<int32_t> ship;
  std::vector<int64_t> qty, price, disc, tax;
  std::vector<char> rf, ls;
  std::string line;
  uint64_t line_no = 0;
  while (std::getline(in, line)) {
    line_no++;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    const char* f[16];
    const char* fe[16];
    int k = 0;
    const char* s = line.c_str();
    const char* start = s;
    for (const char* c = s;; c++) {
      if (*c == '|' || *c == 0) {
        if (k < 16) { f[k] = start; fe[k] = c; k++; }
        start = c + 1;
        if (*c == 0) break;
      }
    }
    if (k == 1 && fe[0] == f[0]) continue;  // empty line
    const std::string where = path + " line " + std::to_string(line_no);
    if (k < 11) throw std::runtime_error("fewer than 11 fields in " + where);
    int64_t v[4];
    for (int c = 0; c < 4; c++)
      if (!parse_decimal(f[4 + c], fe[4 + c], &v[c])) throw std::runtime_error("bad decimal in " + where);
    if (fe[8] - f[8] != 1 || fe[9] - f[9] != 1) throw std::runtime_error("bad return flag or line status in " + where);
    qty.push_back(v[0]);
    price.push_back(v[1]);
    disc.push_back(v[2]);
    tax.push_back(v[3]);
    rf.push_back(*f[8]);
    ls.push_back(*f[9]);
    int32_t d = 0;
    if (!parse_date(std::string(f[10], fe[10]), &d)) throw std::runtime_error("bad ship date in " + where);
    ship.push_back(d);
  }
  const uint64_t n = ship.size();
  if (n == 0) throw std::runtime_error("The lineitem table column cardinality should not be 0");
  auto h = std::make_unique<HostColumns>(ctx, Q1_LAYOUT_REFERENCE, n);
  std::memcpy((void*)h->cols.ship_date, ship.data(), n * 4);
  std::memcpy((void*)h->cols.quantity, qty.data(), n * 8);
  std::memcpy((void*)h->cols.extended_price, price.data(), n * 8);
  std::memcpy((void*)h->cols.discount, disc.data(), n * 8);
  std::memcpy((void*)h->cols.tax, tax.data(), n * 8);
  std::memcpy((void*)h->cols.return_flag, rf.data(), n);
  std::memcpy((void*)h->cols.line_status, ls.data(), n);
  return h;
}

void print_results(const q1_result& r, uint64_t cardinality) {
  std::printf("+---------------------------------------------------- Results ------------------------------------------------------+\n");
  std::printf("|  RF | LS |  sum_quantity        |  sum_base_price      |  sum_disc_price          |  sum_charge              | count      |\n");
  std::printf("+-------------------------------------------------------------------------------------------------------------------+\n");
  uint64_t total = 0;
  auto dec = [](__int128 v, int scale_digits) {
    char buf[64];
    const bool neg = v < 0;
    unsigned __int128 u = neg ? (unsigned __int128)(-v) : (unsigned __int128)v;
    unsigned __int128 pow = 1;
    for (int i = 0; i < scale_digits; i++) pow *= 10;
    const unsigned __int128 ip = u / pow, fp = u % pow;
    std::snprintf(buf, sizeof(buf), "%s%llu.%0*llu", neg ? "-" : "", (unsigned long long)ip, scale_digits,
                  (unsigned long long)fp);
    return std::string(buf);
  };
  for (int g = 0; g < Q1_NUM_GROUPS; g++) {
    const q1_group& x = r.g[g];
    if (x.count == 0) continue;
    const char rf = "ANR"[g >> 1], ls = "FO"[g & 1];
    const __int128 dp = (__int128)x.sum_disc_price_hi << 64 | x.sum_disc_price_lo;
    const __int128 ch = (__int128)x.sum_charge_hi << 64 | x.sum_charge_lo;
    std::printf("| # %c | %c  | %20s | %20s | %24s | %24s | %10lld |\n", rf, ls, dec(x.sum_qty, 2).c_str(),
                dec(x.sum_base_price, 2).c_str(), dec(dp, 4).c_str(), dec(ch, 6).c_str(), (long long)x.count);
    total += (uint64_t)x.count;
  }
  std::printf("+-------------------------------------------------------------------------------------------------------------------+\n");
  std::printf("Total number of elements tuples satisfying the WHERE clause: %llu / %llu = %f\n",
              (unsigned long long)total, (unsigned long long)cardinality, cardinality ? (double)total / cardinality : 0.0);
}

void check(q1_ctx* ctx, q1_status s, const char* what) {
  if (s != Q1_OK) {
    std::fprintf(stderr, "%s failed: %s (%s)\n", what, q1_strerror(s), q1_last_error(ctx));
    std::exit(EXIT_FAILURE);
  }
}

}  // namespace

int main(int argc, char** argv) {
  bool want_device = false;
  const Params p = parse(argc, argv, &want_device);
  if (p.ingest_only) {
    // .tbl -> column cache on the host only (q1_init is never called)
    try {
      auto ref = parse_tbl(nullptr, sf_dir(p) + "/lineitem.tbl");
      if (!ref) throw std::runtime_error("Cannot locate table text file " + sf_dir(p) + "/lineitem.tbl");
      std::printf("CSV read & parsed; table length: %llu records.\n", (unsigned long long)ref->rows);
      write_bin_cache(p, *ref);
      std::printf("Column cache written to %s\n", sf_dir(p).c_str());
    } catch (const std::exception& e) {
      std::fprintf(stderr, "%s\n", e.what());
      return EXIT_FAILURE;
    }
    return 0;
  }
  q1_ctx* ctx = nullptr;
  const q1_status st = q1_init(0, &ctx);
  if (st != Q1_OK) {
    std::fprintf(stderr, "q1_init failed: %s\n", q1_strerror(st));
    return EXIT_FAILURE;
  }
  if (want_device) {
    print_device(ctx);
    q1_shutdown(ctx);
    return 1;  // src/main.cu:266-268 exits after printing
  }
  const q1_placement placement = (q1_placement)placement_code(p.placement);
  q1_launch launch{p.threads_per_block, p.tuples_per_thread, 0, 0};

  // ---- load (untimed, as in the reference) ------------------------------------------------
  std::unique_ptr<HostColumns> ref;
  bool from_generator = false;
  try {
    if (!p.data_dir.empty()) {
      ref = load_bin_cache(ctx, p);
      if (!ref) {
        ref = parse_tbl(ctx, sf_dir(p) + "/lineitem.tbl");
        if (!ref) throw std::runtime_error("Cannot locate table text file " + sf_dir(p) + "/lineitem.tbl");
        std::printf("CSV read & parsed; table length: %llu records.\n", (unsigned long long)ref->rows);
        write_bin_cache(p, *ref);
      }
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return EXIT_FAILURE;
  }
  uint64_t rows = ref ? ref->rows : p.rows;
  std::printf("Lineitem table cardinality for scale factor %f is %llu\n", p.scale_factor, (unsigned long long)rows);
  const q1_layout layout = p.compression ? (p.filter_pushdown ? Q1_LAYOUT_COMPRESSED_PUSHDOWN : Q1_LAYOUT_COMPRESSED)
                                         : Q1_LAYOUT_REFERENCE;
  std::unique_ptr<HostColumns> comp;
  if (p.compression) {
    std::printf("Preprocessing/compressing column data... ");
    std::fflush(stdout);
    comp = std::make_unique<HostColumns>(ctx, Q1_LAYOUT_COMPRESSED, rows);
    if (ref) {
      check(ctx, q1_compress_columns(&ref->cols, rows, &comp->cols, p.cpu_threads), "q1_compress_columns");
    } else {
      check(ctx, q1_generate_host(Q1_LAYOUT_COMPRESSED, p.seed, 0, rows, &comp->cols, p.cpu_threads), "q1_generate_host");
      from_generator = true;
    }
    std::printf("done.\n");
  } else if (!ref) {
    ref = std::make_unique<HostColumns>(ctx, Q1_LAYOUT_REFERENCE, rows);
    check(ctx, q1_generate_host(Q1_LAYOUT_REFERENCE, p.seed, 0, rows, &ref->cols, p.cpu_threads), "q1_generate_host");
    from_generator = true;
  }
  HostColumns& host = p.compression ? *comp : *ref;
  // The narrow kernels run only on rows known to lie in the TPC-H value domain: generated
  // rows do by construction; rows read from a .tbl file or the .bin cache are checked here.
  int in_domain = 1;
  if (!from_generator)
    check(ctx,
          q1_host_columns_in_tpch_domain(p.compression ? Q1_LAYOUT_COMPRESSED : Q1_LAYOUT_REFERENCE, &host.cols, rows,
                                         p.cpu_threads, &in_domain),
          "q1_host_columns_in_tpch_domain");
  if (in_domain) launch.flags |= Q1_LAUNCH_ASSUME_TPCH_DOMAIN;
  if (!in_domain) std::printf("Values outside the TPC-H domain: the 64-bit kernels run.\n");
  std::unique_ptr<HostColumns> pushdown;
  if (p.filter_pushdown) {
    pushdown = std::make_unique<HostColumns>(ctx, Q1_LAYOUT_COMPRESSED_PUSHDOWN, rows);
    q1_columns& pc = pushdown->cols;
    // share the compressed value columns; only the bitmap is new
    uint64_t b[8];
    q1_layout_column_bytes(Q1_LAYOUT_COMPRESSED_PUSHDOWN, rows, b);
    std::memcpy((void*)pc.quantity, comp->cols.quantity, b[1]);
    std::memcpy((void*)pc.extended_price, comp->cols.extended_price, b[2]);
    std::memcpy((void*)pc.discount, comp->cols.discount, b[3]);
    std::memcpy((void*)pc.tax, comp->cols.tax, b[4]);
    std::memcpy((void*)pc.return_flag, comp->cols.return_flag, b[5]);
    std::memcpy((void*)pc.line_status, comp->cols.line_status, b[6]);
  }
  q1_table* resident = nullptr;
  if (p.device_resident) {
    check(ctx, q1_table_upload(ctx, p.compression ? Q1_LAYOUT_COMPRESSED : Q1_LAYOUT_REFERENCE, &host.cols, rows,
                               &resident),
          "q1_table_upload");
    q1_timing t;
    q1_last_timing(ctx, &t);
    std::printf("Uploaded %llu bytes in %.3f ms (%.2f GB/s)\n", (unsigned long long)t.bytes_read, t.h2d_ms,
                t.bytes_read / (t.h2d_ms * 1e-3) / 1e9);
  }

  // Pushdown bitmaps: the host's for streamed rows and for the CPU share of a split, the
  // device's for resident rows.  Built in every timed run (as the reference's producer
  // does, src/main.cu:362-389), or once before the runs with --filter-once.
  const bool host_filter = p.filter_pushdown && (!resident || p.coprocessing);
  const bool device_filter = p.filter_pushdown && resident;
  auto build_host_filter = [&] {
    check(ctx, q1_cpu_build_filter((const uint16_t*)comp->cols.ship_date, rows, p.cutoff,
                                   (uint32_t*)pushdown->cols.filter, p.cpu_threads),
          "q1_cpu_build_filter");
  };
  q1_table* resident_filter = nullptr;
  if (p.filter_pushdown && p.filter_once) {
    const auto fb = std::chrono::steady_clock::now();
    if (host_filter) build_host_filter();
    if (device_filter) check(ctx, q1_build_filter(ctx, resident, p.cutoff, &resident_filter), "q1_build_filter");
    std::printf("Filter bitmap built once in %.3f ms\n",
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - fb).count());
  }

  std::ofstream results_file("results.csv");
  for (int run = 0; run < p.runs; run++) {
    std::printf("Executing TPC-H Query 1, run %d of %d... ", run + 1, p.runs);
    std::fflush(stdout);
    const auto start = std::chrono::steady_clock::now();
    q1_result result{};
    const uint64_t cpu_start = p.coprocessing ? (rows - (uint64_t)(rows * p.cpu_fraction)) / p.tuples_per_launch *
                                                    p.tuples_per_launch
                                              : rows;
    // the host evaluates the WHERE clause into the bitmap (src/main.cu:362-389), in the timed run
    if (host_filter && !p.filter_once) build_host_filter();
    const q1_columns& run_cols = p.filter_pushdown ? pushdown->cols : host.cols;
    q1_result cpu_part{};
    std::thread cpu_thread;
    q1_status cpu_status = Q1_OK;
    if (cpu_start < rows) {
      cpu_thread = std::thread([&] {
        cpu_status = q1_cpu_run(layout, &run_cols, cpu_start, rows - cpu_start, p.cutoff, p.cpu_threads, &cpu_part);
      });
    }
    if (resident) {
      q1_table* t = resident;
      q1_table* pd = nullptr;
      if (resident_filter) {
        t = resident_filter;
      } else if (p.filter_pushdown) {
        // the device evaluates the WHERE clause into the bitmap, in the timed run
        check(ctx, q1_build_filter(ctx, resident, p.cutoff, &pd), "q1_build_filter");
        t = pd;
      }
      check(ctx, q1_run(ctx, t, 0, cpu_start, p.cutoff, placement, &launch, &result), "q1_run");
      if (pd) q1_table_free(pd);
    } else if (cpu_start > 0) {
      check(ctx, q1_run_streaming(ctx, layout, &run_cols, cpu_start, p.cutoff, placement, &launch, p.tuples_per_launch,
                                  p.streams, &result),
            "q1_run_streaming");
    }
    if (cpu_thread.joinable()) cpu_thread.join();
    check(ctx, cpu_status, "q1_cpu_run");
    if (cpu_start < rows) q1_result_merge(&result, &cpu_part, &result);
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    q1_timing tm;
    q1_last_timing(ctx, &tm);
    std::printf("done. (%.3f ms wall, %.3f ms device)\n", sec * 1e3, tm.kernel_ms);
    results_file << sec << '\n';
    if (p.print_results) print_results(result, rows);
  }
  if (resident_filter) q1_table_free(resident_filter);  // before the table whose columns it shares
  if (resident) q1_table_free(resident);
  q1_shutdown(ctx);
  return 0;
}
