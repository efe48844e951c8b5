// Input/output formats of the `kgwas` drop-in (host side, no GPU code).
//
// Formats and their quirks follow the reference exactly:
//   ACCESSIONS  one id per line, empty lines skipped, one trailing '\r'
//               stripped                       (src/kgwas_utils.hpp:51-61)
//   PHENOTYPE   whitespace-separated "id value" pairs read until the first
//               pair that does not parse; a repeated id keeps the last value
//                                              (src/kgwas_utils.hpp:63-70)
//   overlap     accession positions with a phenotype, in accession order
//                                              (src/kgwas_utils.hpp:291-314)
//   U.bin       int32 n, int32 r, n*r float64 column-major; n must equal the
//               overlap size                   (src/kgwas_utils.hpp:135-157)
//   batch dir   bits/col_%06d.txt (mmap) or .txt.gz, one byte per k-mer,
//               value = byte & 1; kmers.txt.gz with exactly N lines
//                                              (src/kgwas_utils.hpp:72-127,
//                                               166-282, 317-332)
//   outputs     "{out}", "{out with .2d before the last dot}", "{out}.count"
//                                              (src/kgwas.cpp:225-304)
// The batch is packed here into k-mer-major rows of uint64 words (accession a
// = bit a%64 of word a/64), the layout the scan kernels stream.
#pragma once
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <emmintrin.h>
#include <zlib.h>

#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

namespace kgio {

using std::string;
using std::vector;

struct Clock {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double sec() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

[[noreturn]] inline void die(const string& m) {
  std::cerr << "[ERROR] " << m << "\n";
  std::exit(1);
}

inline void warn(const string& m) { std::cerr << "[WARN] " << m << "\n"; }

// Integer environment value, floor 1; unset/empty/invalid -> default.
inline int env_int(const char* k, int dflt) {
  const char* v = std::getenv(k);
  if (!v || !*v) return dflt;
  try {
    int x = std::stoi(v);
    return x < 1 ? 1 : x;
  } catch (...) {
    return dflt;
  }
}

inline bool is_file(const string& p) {
  struct stat st {};
  return ::stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

inline bool has_suffix(const string& s, const string& suf) {
  return s.size() >= suf.size() && s.compare(s.size() - suf.size(), suf.size(), suf) == 0;
}

inline vector<string> read_accessions(const string& path) {
  std::ifstream in(path);
  if (!in) die("open: " + path);
  vector<string> out;
  string line;
  while (std::getline(in, line)) {
    if (line.empty()) continue;
    if (line.back() == '\r') line.pop_back();
    out.push_back(line);
  }
  return out;
}

inline std::unordered_map<string, double> read_phenotype(const string& path) {
  std::ifstream in(path);
  if (!in) die("open: " + path);
  std::unordered_map<string, double> m;
  string id;
  double v;
  while (in >> id >> v) m[id] = v;
  return m;
}

struct Overlap {
  vector<string> accessions;
  vector<int32_t> idxs;  // accession positions with a phenotype
  vector<double> y;      // aligned phenotype
};

inline Overlap align_samples(const string& acc_path, const string& pheno_path) {
  Clock t;
  Overlap ov;
  ov.accessions = read_accessions(acc_path);
  if (ov.accessions.empty()) die("samples empty");
  auto ph = read_phenotype(pheno_path);
  for (size_t j = 0; j < ov.accessions.size(); ++j) {
    auto it = ph.find(ov.accessions[j]);
    if (it != ph.end()) {
      ov.idxs.push_back((int32_t)j);
      ov.y.push_back(it->second);
    }
  }
  if (ov.idxs.empty()) die("no phenotype overlap");
  std::cerr << "[INFO] n=" << ov.y.size() << " (overlap)  (" << t.sec() << " s)\n";
  return ov;
}

struct UBin {
  int n = 0, r = 0;
  vector<double> U;  // column-major n x r
};

inline UBin read_ubin(const string& path, int n_overlap) {
  std::ifstream in(path, std::ios::binary);
  if (!in) die("open: " + path);
  UBin u;
  int32_t hdr[2] = {0, 0};
  in.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
  if (!in) die("short U.bin header");
  u.n = hdr[0];
  u.r = hdr[1];
  if (u.n != n_overlap)
    die("U.bin n mismatch: file n=" + std::to_string(u.n) +
        " vs overlap n=" + std::to_string(n_overlap));
  if (u.r < 0) die("U.bin r < 0");
  u.U.resize((size_t)u.n * (size_t)u.r);
  in.read(reinterpret_cast<char*>(u.U.data()), (std::streamsize)(sizeof(double) * u.U.size()));
  if (!in) die("short U.bin body");
  std::cerr << "[INFO] loaded U (n=" << u.n << ", r=" << u.r << ")\n";
  return u;
}

// ---- batch directory -------------------------------------------------------

inline string col_name(const string& bits_dir, int acc) {
  char base[64];
  std::snprintf(base, sizeof(base), "col_%06d.txt", acc);
  return bits_dir + "/" + base;
}

// Bytes before the first '\n' (or the whole file) of a column file.
inline size_t column_length(const string& bits_dir, int acc) {
  const string txt = col_name(bits_dir, acc), gz = txt + ".gz";
  if (is_file(txt)) {
    int fd = ::open(txt.c_str(), O_RDONLY);
    if (fd < 0) die("open: " + txt);
    struct stat st {};
    if (fstat(fd, &st) != 0) {
      ::close(fd);
      die("fstat: " + txt);
    }
    const size_t sz = (size_t)st.st_size;
    if (sz == 0) {
      ::close(fd);
      die("empty: " + txt);
    }
    void* p = mmap(nullptr, sz, PROT_READ, MAP_PRIVATE, fd, 0);
    if (p == MAP_FAILED) {
      ::close(fd);
      die("mmap: " + txt);
    }
    const void* nl = memchr(p, '\n', sz);
    const size_t N = nl ? (size_t)((const char*)nl - (const char*)p) : sz;
    munmap(p, sz);
    ::close(fd);
    if (N == 0) die("no bytes before newline: " + txt);
    std::cerr << "[INFO] inferred N=" << N << " rows [from .txt]\n";
    return N;
  }
  if (is_file(gz)) {
    gzFile f = gzopen(gz.c_str(), "rb");
    if (!f) die("gzopen: " + gz);
    gzbuffer(f, 1 << 20);
    vector<unsigned char> buf(1 << 20);
    size_t total = 0;
    for (;;) {
      const int got = gzread(f, buf.data(), (unsigned)buf.size());
      if (got < 0) {
        int err = 0;
        const char* msg = gzerror(f, &err);
        gzclose(f);
        die(string("gzread error: ") + (msg ? msg : ""));
      }
      if (got == 0) break;
      const void* nl = memchr(buf.data(), '\n', (size_t)got);
      if (nl) {
        total += (size_t)((const unsigned char*)nl - buf.data());
        gzclose(f);
        if (total == 0) die("N=0");
        std::cerr << "[INFO] inferred N=" << total << " rows [from .txt.gz]\n";
        return total;
      }
      total += (size_t)got;
    }
    gzclose(f);
    if (total == 0) die("empty gz column: " + gz);
    std::cerr << "[INFO] inferred N=" << total << " rows [from .txt.gz]\n";
    return total;
  }
  die("missing column: " + txt + " (and " + gz + ")");
}

// kmers.txt(.gz): exactly N lines (gz: split at '\n' only; plain: strip '\r').
inline vector<string> read_kmers(const string& path, size_t N) {
  vector<string> km;
  km.reserve(N);
  if (has_suffix(path, ".gz")) {
    gzFile f = gzopen(path.c_str(), "rb");
    if (!f) die("gzopen: " + path);
    gzbuffer(f, 1 << 20);
    vector<char> buf(1 << 15);
    string line;
    while (km.size() < N) {
      char* r = gzgets(f, buf.data(), (int)buf.size());
      if (!r) {
        int err = 0;
        const char* msg = gzerror(f, &err);
        if (err != Z_OK && err != Z_STREAM_END) {
          gzclose(f);
          die(string("gzgets error: ") + (msg ? msg : ""));
        }
        break;
      }
      char* nl = strchr(r, '\n');
      if (nl) {
        *nl = '\0';
        line.append(r);
        km.push_back(std::move(line));
        line.clear();
      } else {
        line.append(r);
      }
    }
    gzclose(f);
  } else {
    std::ifstream in(path);
    if (!in) die("open: " + path);
    string s;
    while (km.size() < N && std::getline(in, s)) {
      if (!s.empty() && s.back() == '\r') s.pop_back();
      km.push_back(std::move(s));
    }
  }
  if (km.size() != N)
    die("kmers lines != N (got " + std::to_string(km.size()) + ", expected " +
        std::to_string(N) + ")");
  return km;
}

// One phenotyped sample's column: N bytes (mmap of .txt or decompressed .gz).
struct Column {
  const unsigned char* data = nullptr;
  size_t map_len = 0;
  vector<unsigned char> owned;
  ~Column() {
    if (map_len) munmap(const_cast<unsigned char*>(data), map_len);
  }
};

inline void open_column(const string& bits_dir, int acc, size_t N, Column& col) {
  const string txt = col_name(bits_dir, acc), gz = txt + ".gz";
  if (is_file(txt)) {
    int fd = ::open(txt.c_str(), O_RDONLY);
    if (fd < 0) die("open: " + txt);
    struct stat st {};
    if (fstat(fd, &st) != 0) {
      ::close(fd);
      die("fstat: " + txt);
    }
    const size_t sz = (size_t)st.st_size;
    if (sz < N) {
      ::close(fd);
      die("short column: " + txt + " (size " + std::to_string(sz) + " < N " + std::to_string(N) +
          ")");
    }
    // page tables filled here, by the thread that opens the column
    void* p = mmap(nullptr, sz, PROT_READ, MAP_PRIVATE | MAP_POPULATE, fd, 0);
    ::close(fd);
    if (p == MAP_FAILED) die("mmap: " + txt);
    col.data = (const unsigned char*)p;
    col.map_len = sz;
    return;
  }
  if (is_file(gz)) {
    gzFile f = gzopen(gz.c_str(), "rb");
    if (!f) die("gzopen: " + gz);
    gzbuffer(f, 1 << 20);
    col.owned.resize(N);
    size_t got = 0;
    while (got < N) {
      const unsigned want = (unsigned)std::min<size_t>(N - got, 1u << 20);
      const int r = gzread(f, col.owned.data() + got, want);
      if (r <= 0) {
        int err = 0;
        const char* msg = gzerror(f, &err);
        gzclose(f);
        if (r == 0 || err == Z_STREAM_END || err == Z_OK)
          die("short gz column: " + gz + " (got " + std::to_string(got) + " < N " +
              std::to_string(N) + ")");
        die(string("gzread error: ") + (msg ? msg : ""));
      }
      got += (size_t)r;
    }
    unsigned char extra = 0;
    const int r = gzread(f, &extra, 1);
    gzclose(f);
    if (r == 1 && extra != '\n')
      die("gz column longer than expected (N+ extra non-newline): " + gz);
    col.data = col.owned.data();
    return;
  }
  die("missing column: " + txt + " (and " + gz + ")");
}

// 16 accessions x 16 k-mers of column bytes -> one 16-bit mask per k-mer (bit q
// = byte & 1 of accession q): a 16 x 16 byte transpose by unpacks, then
// movemask of each k-mer's 16 bytes shifted so that bit 0 is the sign bit.
inline void pack_tile16(const unsigned char* const* col, size_t i, uint16_t* out,
                        size_t out_stride) {
  __m128i a[16], b[16];
  for (int q = 0; q < 16; ++q) a[q] = _mm_loadu_si128((const __m128i*)(col[q] + i));
  // a[q] byte k = (accession q, k-mer k).  Each stage interleaves pairs of
  // vectors, doubling the run of accessions held together per k-mer.
  for (int j = 0; j < 8; ++j) {  // 2 accessions x 8 k-mers per vector
    b[j] = _mm_unpacklo_epi8(a[2 * j], a[2 * j + 1]);
    b[j + 8] = _mm_unpackhi_epi8(a[2 * j], a[2 * j + 1]);
  }
  for (int j = 0; j < 4; ++j) {  // 4 accessions x 4 k-mers
    a[j] = _mm_unpacklo_epi16(b[2 * j], b[2 * j + 1]);
    a[j + 4] = _mm_unpackhi_epi16(b[2 * j], b[2 * j + 1]);
    a[j + 8] = _mm_unpacklo_epi16(b[8 + 2 * j], b[8 + 2 * j + 1]);
    a[j + 12] = _mm_unpackhi_epi16(b[8 + 2 * j], b[8 + 2 * j + 1]);
  }
  for (int h = 0; h < 4; ++h)    // 8 accessions x 2 k-mers
    for (int j = 0; j < 2; ++j) {
      b[4 * h + j] = _mm_unpacklo_epi32(a[4 * h + 2 * j], a[4 * h + 2 * j + 1]);
      b[4 * h + j + 2] = _mm_unpackhi_epi32(a[4 * h + 2 * j], a[4 * h + 2 * j + 1]);
    }
  for (int h = 0; h < 8; ++h) {  // 16 accessions x 1 k-mer
    a[2 * h] = _mm_unpacklo_epi64(b[2 * h], b[2 * h + 1]);
    a[2 * h + 1] = _mm_unpackhi_epi64(b[2 * h], b[2 * h + 1]);
  }
  for (int k = 0; k < 16; ++k)
    out[(size_t)k * out_stride] = (uint16_t)_mm_movemask_epi8(_mm_slli_epi64(a[k], 7));
}

// Packs the phenotyped samples' columns into k-mer-major rows:
// rows[i * row_words + a/64] bit a%64 = byte_i(col of accession a) & 1 (0 for
// accessions without a phenotype).  Tiles of 16 accessions x 16 k-mers are
// transposed in SSE2 registers and stored as 16-bit row pieces; an accession
// without a phenotype reads a zero column.
inline vector<uint64_t> pack_batch(const string& bits_dir, const vector<int32_t>& idxs, size_t N,
                                   int64_t row_words, int threads) {
  Clock t;
  const int n = (int)idxs.size();
  vector<Column> cols(n);
  std::atomic<int> done{0};
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
  for (int k = 0; k < n; ++k) {
    open_column(bits_dir, idxs[k], N, cols[k]);
    const int d = ++done;
    if (d % 64 == 0) {
#pragma omp critical
      std::cerr << "\r[LOAD] " << d << "/" << n << std::flush;
    }
  }
  const double t_open = t.sec();
  // column of every accession up to the last phenotyped one, in groups of 16
  const int n_acc = n ? idxs.back() + 1 : 0;
  const int ngroups = (n_acc + 15) / 16;
  const int64_t BK = 1024;  // k-mers per tile row (a multiple of 16)
  vector<unsigned char> zeros((size_t)BK, 0);
  vector<const unsigned char*> acc_col((size_t)ngroups * 16, nullptr);
  for (int k = 0; k < n; ++k) acc_col[(size_t)idxs[k]] = cols[(size_t)k].data;
  vector<uint64_t> rows((size_t)N * (size_t)row_words, 0ull);
  const size_t stride16 = (size_t)row_words * 4;  // uint16 pieces per row
  const int64_t nfull = (int64_t)(N / 16) * 16;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
  for (int64_t i0 = 0; i0 < nfull; i0 += BK) {
    const int64_t i1 = std::min<int64_t>(i0 + BK, nfull);
    const unsigned char* cp[16];
    for (int g = 0; g < ngroups; ++g) {
      bool any = false;
      for (int q = 0; q < 16; ++q) {
        const unsigned char* c = acc_col[(size_t)g * 16 + q];
        any |= c != nullptr;
        cp[q] = c ? c + i0 : zeros.data();
      }
      if (!any) continue;
      uint16_t* out = reinterpret_cast<uint16_t*>(rows.data()) + (size_t)i0 * stride16 + g;
      for (int64_t i = 0; i < i1 - i0; i += 16)
        pack_tile16(cp, (size_t)i, out + (size_t)i * stride16, stride16);
    }
  }
  for (int k = 0; k < n; ++k) {  // k-mers past the last full tile of 16
    const int a = idxs[k];
    const unsigned char* c = cols[k].data;
    for (int64_t i = nfull; i < (int64_t)N; ++i)
      rows[(size_t)i * row_words + (a >> 6)] |= (uint64_t)(c[i] & 1u) << (a & 63);
  }
  std::cerr << "\r[LOAD] " << n << "/" << n << "  (" << t.sec() << " s: open " << t_open
            << " s, pack " << t.sec() - t_open << " s; " << N << " k-mers x " << row_words
            << " words)\n";
  return rows;
}

// ".2d" inserted before the last '.' of the path (src/kgwas.cpp:273-278).
inline string path_2d(const string& out) {
  const size_t dot = out.rfind('.');
  if (dot != string::npos) return out.substr(0, dot) + ".2d" + out.substr(dot);
  return out + ".2d";
}

// "{basename without extension}\t{passed}\n" into out + ".count"
// (src/kgwas.cpp:292-304).
inline void write_count(const string& out, int64_t passed) {
  const string cf = out + ".count";
  string shard = out;
  const size_t sl = shard.find_last_of("/\\");
  if (sl != string::npos) shard = shard.substr(sl + 1);
  const size_t dot = shard.find_last_of('.');
  if (dot != string::npos) shard = shard.substr(0, dot);
  std::ofstream co(cf);
  if (co.is_open()) {
    co << shard << "\t" << passed << "\n";
    std::cerr << "[INFO] Count written to " << cf << "\n";
  }
}

}  // namespace kgio
