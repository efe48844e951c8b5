// kgwas -- drop-in for SuperGWAS's 1D k-mer scan binary, computing on one
// MI355X through libkgwas_hip (include/kgwas_hip.h).
//
//   kgwas samples.txt phenotype.txt INPUT U.bin output.tsv [log10_thresh]
//         [--mode cols|rows] [--log10-thresh-2d X]
//
// Command line, environment (LOAD_THREADS, BLAS_THREADS, EMIT_THREADS,
// CORR_THRESHOLD, LOG10_THRESH_2D_INPUT), outputs and error behaviour
// ("[ERROR] msg" + exit 1) follow src/kgwas.cpp:26-126, 168-307 (cols) and
// 343-522 (rows).  Additional environment: KG_DEVICE (HIP device, default 0),
// KG_PRECISION (mixed | fp32, default mixed), KG_BATCH_KMERS (rows-mode batch),
// KG_PACK_ONLY (cols mode: write the packed rows, uint64 [N][row_words], to
// this path and stop before the scan), KG_DEDUP (1: scan each run of equal
// adjacent rows once, kg_set_dedup; same outputs).
//
// Numerics: every reported k-mer's effect/T/p come from FP64 values computed
// on the GPU and the reference's compute_hit on the host, so cols mode prints
// the FP64 (rows-mode) statistic rather than the reference's FP32 one; with
// KG_PRECISION=fp32 the GPU reproduces the FP32 cols-mode arithmetic.
#include <omp.h>
#include <sys/stat.h>
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <sstream>

#include "kgwas_hip.h"
#include "kgwas_io.hpp"

using namespace kgio;

namespace {

struct Config {
  string samples, pheno, input, ubin, out;
  double log10_thresh = 6.0;
  double log10_thresh_2d = 15.0;
  double corr = 0.2;
  int load_threads = 24, blas_threads = 112, emit_threads = 112;
  string mode;
};

double parse_double(const string& s, const string& what) {
  try {
    return std::stod(s);
  } catch (...) {
    die("invalid " + what + ": " + s);
  }
}

Config parse_args(int argc, char** argv) {  // src/kgwas.cpp:67-126
  if (argc < 6) {
    std::cerr << "Usage: " << argv[0]
              << " samples.txt phenotype.txt INPUT U.bin output.tsv [log10_thresh] [--mode "
                 "cols|rows]\n\n"
              << "  INPUT: directory (batch with bits/ + kmers.txt.gz) for COLS mode,\n"
              << "         or a .tsv[.gz] file for ROWS mode.\n"
              << "  Auto-detects mode from INPUT type unless --mode is specified.\n\n"
              << "  Env: LOAD_THREADS, BLAS_THREADS, EMIT_THREADS, CORR_THRESHOLD, "
                 "LOG10_THRESH_2D_INPUT, KG_DEVICE, KG_PRECISION\n";
    std::exit(1);
  }
  Config c;
  c.samples = argv[1];
  c.pheno = argv[2];
  c.input = argv[3];
  c.ubin = argv[4];
  c.out = argv[5];
  for (int i = 6; i < argc; ++i) {
    const string a = argv[i];
    if (a == "--mode" && i + 1 < argc) {
      c.mode = argv[++i];
    } else if (a == "--log10-thresh-2d" && i + 1 < argc) {
      c.log10_thresh_2d = parse_double(argv[++i], "--log10-thresh-2d");
    } else {
      try {
        c.log10_thresh = std::stod(a);
      } catch (...) {
        die("invalid argument: " + a);
      }
      if (c.log10_thresh < 0.0) die("threshold must be >= 0");
    }
  }
  c.load_threads = env_int("LOAD_THREADS", c.load_threads);
  c.blas_threads = env_int("BLAS_THREADS", c.blas_threads);
  c.emit_threads = env_int("EMIT_THREADS", c.emit_threads);
  if (const char* v = std::getenv("CORR_THRESHOLD")) c.corr = parse_double(v, "CORR_THRESHOLD");
  if (const char* v = std::getenv("LOG10_THRESH_2D_INPUT"))
    c.log10_thresh_2d = parse_double(v, "LOG10_THRESH_2D_INPUT");
  if (c.mode.empty()) {
    struct stat st {};
    if (::stat(c.input.c_str(), &st) != 0) die("cannot stat INPUT: " + c.input);
    c.mode = S_ISDIR(st.st_mode) ? "cols" : "rows";
    std::cerr << "[INFO] auto-detected mode: " << c.mode << "\n";
  }
  if (c.mode != "cols" && c.mode != "rows")
    die("--mode must be 'cols' or 'rows', got: " + c.mode);
  return c;
}

// Device context + null model; any library failure ends the run like die().
struct Scanner {
  kg_ctx* ctx = nullptr;
  int t = 1;
  Scanner(int n_acc, const Overlap& ov, const UBin& u) {
    kg_opts o{};
    const char* dv = std::getenv("KG_DEVICE");
    o.device = dv ? std::atoi(dv) : 0;
    const char* pv = std::getenv("KG_PRECISION");
    const string prec = pv ? pv : "mixed";
    if (prec == "mixed") o.precision = KG_PREC_MIXED;
    else if (prec == "fp32") o.precision = KG_PREC_FP32;
    else die("KG_PRECISION must be 'mixed' or 'fp32', got: " + prec);
    int rc = kg_create(&ctx, &o);
    if (rc != KG_OK) die(string("GPU init failed: ") + (ctx ? kg_last_error(ctx) : "kg_create"));
    rc = kg_set_model(ctx, n_acc, ov.idxs.data(), (int32_t)ov.idxs.size(),
                      u.U.empty() ? nullptr : u.U.data(), u.r, ov.y.data(), 1);
    if (rc != KG_OK) die(string("model set-up failed: ") + kg_last_error(ctx));
    if (const char* dd = std::getenv("KG_DEDUP")) {  // batches sorted by bits: runs of equal rows
      const string v = dd;
      if (v != "0" && v != "1") die("KG_DEDUP must be 0 or 1, got: " + v);
      if (kg_set_dedup(ctx, v == "1" ? 1 : 0) != KG_OK) die(kg_last_error(ctx));
    }
  }
  ~Scanner() { kg_destroy(ctx); }
  vector<kg_hit> scan(const vector<uint64_t>& rows, int64_t m, int64_t rw,
                      const kg_scan_params& prm, int64_t* passed) {
    vector<kg_hit> hits(1024);
    for (;;) {
      int64_t nout = 0;
      const int rc =
          kg_scan(ctx, rows.data(), m, rw, &prm, hits.data(), (int64_t)hits.size(), &nout, passed);
      if (rc == KG_ERR_CAPACITY) {
        hits.resize((size_t)nout);
        continue;
      }
      if (rc != KG_OK) die(string("scan failed: ") + kg_last_error(ctx));
      hits.resize((size_t)nout);
      return hits;
    }
  }
};

int64_t row_words_for(int n_acc) {
  const int64_t w = (n_acc + 63) / 64;
  return (w + 7) / 8 * 8;
}

void warn_yyM(const Overlap& ov, const UBin& u) {  // src/kgwas_utils.hpp:151-153
  const int n = (int)ov.y.size();
  double yy = 0.0, uu = 0.0;
  for (double v : ov.y) yy += v * v;
  for (int k = 0; k < u.r; ++k) {
    double s = 0.0;
    for (int j = 0; j < n; ++j) s += u.U[(size_t)k * n + j] * ov.y[j];
    uu += s * s;
  }
  if (!(yy - uu > 1e-12)) warn("yyM ~ 0");
}

// ---------------------------------------------------------------------------
void run_cols(const Config& cfg) {
  Clock tall;
  Overlap ov = align_samples(cfg.samples, cfg.pheno);
  UBin u = read_ubin(cfg.ubin, (int)ov.y.size());
  warn_yyM(ov, u);
  const string bits_dir = cfg.input + "/bits";
  const string kmers_gz = cfg.input + "/kmers.txt.gz";
  const size_t N = column_length(bits_dir, ov.idxs.front());
  vector<string> kmers = read_kmers(kmers_gz, N);
  const int n_acc = (int)ov.accessions.size();
  const int64_t rw = row_words_for(n_acc);
  vector<uint64_t> rows = pack_batch(bits_dir, ov.idxs, N, rw, cfg.load_threads);
  if (const char* dump = std::getenv("KG_PACK_ONLY")) {  // diagnostic: packed rows, no scan
    std::ofstream out(dump, std::ios::binary);
    out.write(reinterpret_cast<const char*>(rows.data()),
              (std::streamsize)(sizeof(uint64_t) * rows.size()));
    if (!out) die("write: " + string(dump));
    return;
  }

  Clock tscan;
  Scanner sc(n_acc, ov, u);
  const kg_scan_params prm{cfg.corr, cfg.log10_thresh, cfg.log10_thresh_2d};
  int64_t passed = 0;
  vector<kg_hit> hits = sc.scan(rows, (int64_t)N, rw, prm, &passed);
  std::cerr << "[INFO] CORR_THRESHOLD=" << cfg.corr << "  total kmers: " << N
            << ", passed correlation filter: " << passed << "\n";
  // the reference's stage lines (src/kgwas.cpp:217, 164): device time of the
  // pre-filter screen, then of the projections and the FP64 refinement
  kg_timing tm{};
  if (kg_last_timing(sc.ctx, &tm) != KG_OK) die(string("timing: ") + kg_last_error(sc.ctx));
  std::cerr << "[INFO] Correlation filter (" << tm.screen_ms * 1e-3 << " s)\n";
  std::cerr << "[INFO] LA (" << (tm.bulk_ms + tm.refine_ms) * 1e-3
            << " s, BLAS_THREADS=" << cfg.blas_threads << ")\n";
  std::cerr << "[INFO] GPU scan (" << tscan.sec() << " s, copies " << tm.copy_ms * 1e-3
            << " s)\n";

  // emit in k-mer order (src/kgwas.cpp:239-266)
  Clock temit;
  const bool do_2d = cfg.log10_thresh_2d > 0.0;
  const int n = (int)ov.idxs.size();
  std::ostringstream o1, o2;
  o1.setf(std::ios::fmtflags(0), std::ios::floatfield);
  o2.setf(std::ios::fmtflags(0), std::ios::floatfield);
  string bits(n, '0');
  for (const kg_hit& h : hits) {
    o1 << kmers[(size_t)h.idx] << '\t' << h.effect << '\t' << h.pval << '\n';
    if (do_2d && (h.flags & KG_HIT_2D)) {
      const uint64_t* row = rows.data() + (size_t)h.idx * rw;
      for (int j = 0; j < n; ++j) {
        const int a = ov.idxs[j];
        bits[j] = ((row[a >> 6] >> (a & 63)) & 1ull) ? '1' : '0';
      }
      o2 << kmers[(size_t)h.idx] << '\t' << bits << '\t' << h.effect << '\t' << h.pval << '\n';
    }
  }
  {
    std::ofstream out(cfg.out, std::ios::binary);
    if (!out) die("open out: " + cfg.out);
    const string s = o1.str();
    out.write(s.data(), (std::streamsize)s.size());
    out.close();
    if (!out) die("write out failed");
  }
  std::cerr << "[INFO] wrote 1D hits at -log10(p) >= " << cfg.log10_thresh << " ("
            << temit.sec() << " s, EMIT_THREADS=" << cfg.emit_threads << ", " << hits.size()
            << " hits)\n";
  if (do_2d) {
    const string p2 = path_2d(cfg.out);
    std::ofstream out2(p2, std::ios::binary);
    if (!out2) die("open 2D out: " + p2);
    const string s = o2.str();
    out2.write(s.data(), (std::streamsize)s.size());
    out2.close();
    if (!out2) die("write 2D out failed");
    std::cerr << "[INFO] wrote 2D hits at -log10(p) >= " << cfg.log10_thresh_2d << " to " << p2
              << "\n";
  }
  write_count(cfg.out, passed);
  std::cerr << "[TIME] total " << tall.sec() << " s\n";
}

// ---------------------------------------------------------------------------
// Rows mode: INPUT is a .tsv[.gz] of "kmer<TAB|SPACE>bitstring" lines
// (src/kgwas.cpp:343-522).  x(j) = (bits[idxs[j]] == '1'); batches of lines
// are packed and scanned on the GPU; output keeps line order.
bool next_line(gzFile gz, std::ifstream& in, bool use_gz, string& line) {
  line.clear();
  if (use_gz) {
    char buf[1 << 16];
    bool any = false;
    for (;;) {
      char* r = gzgets(gz, buf, sizeof(buf));
      if (!r) return any;
      any = true;
      size_t len = strlen(r);
      if (len == 0) continue;
      if (r[len - 1] == '\n') {
        size_t e = len - 1;
        if (e > 0 && r[e - 1] == '\r') --e;
        line.append(r, e);
        return true;
      }
      line.append(r, len);
    }
  }
  return (bool)std::getline(in, line);
}

void run_rows(const Config& cfg) {
  Clock tall;
  Overlap ov = align_samples(cfg.samples, cfg.pheno);
  UBin u = read_ubin(cfg.ubin, (int)ov.y.size());
  warn_yyM(ov, u);
  const int n_acc = (int)ov.accessions.size();
  const int n = (int)ov.idxs.size();
  const int64_t rw = row_words_for(n_acc);
  const bool use_gz = has_suffix(cfg.input, ".gz");
  gzFile gz = nullptr;
  std::ifstream plain;
  if (use_gz) {
    gz = gzopen(cfg.input.c_str(), "rb");
    if (!gz) die("Cannot open gz file: " + cfg.input);
    gzbuffer(gz, 1u << 20);
  } else {
    plain.open(cfg.input);
    if (!plain) die("Cannot open kmer file: " + cfg.input);
  }
  std::ofstream out(cfg.out);
  if (!out) die("Cannot open output: " + cfg.out);
  const bool do_2d = cfg.log10_thresh_2d > 0.0;
  std::ofstream out2;
  if (do_2d) {
    out2.open(path_2d(cfg.out));
    if (!out2) die("Cannot open 2D output: " + path_2d(cfg.out));
  }
  Scanner sc(n_acc, ov, u);
  const kg_scan_params prm{cfg.corr, cfg.log10_thresh, cfg.log10_thresh_2d};
  const int64_t B = std::max(1, env_int("KG_BATCH_KMERS", 1 << 20));
  int64_t lines = 0, passed_total = 0;
  vector<string> kmer, bitstr;
  vector<uint64_t> rows;
  string line;
  bool eof = false;
  while (!eof) {
    kmer.clear();
    bitstr.clear();
    while ((int64_t)kmer.size() < B) {
      if (!next_line(gz, plain, use_gz, line)) {
        eof = true;
        break;
      }
      if (line.empty()) continue;
      ++lines;
      size_t tab = line.find('\t');
      if (tab == string::npos) {
        tab = line.find(' ');
        if (tab == string::npos) continue;
      }
      kmer.push_back(line.substr(0, tab));
      bitstr.push_back(line.substr(tab + 1));
    }
    const int64_t m = (int64_t)kmer.size();
    if (m == 0) continue;
    rows.assign((size_t)m * rw, 0ull);
    for (int64_t i = 0; i < m; ++i) {
      const string& b = bitstr[(size_t)i];
      uint64_t* row = rows.data() + (size_t)i * rw;
      for (int j = 0; j < n; ++j) {
        const int a = ov.idxs[j];
        if ((size_t)a < b.size() && b[(size_t)a] == '1') row[a >> 6] |= 1ull << (a & 63);
      }
    }
    int64_t passed = 0;
    vector<kg_hit> hits = sc.scan(rows, m, rw, prm, &passed);
    passed_total += passed;
    string ob(n, '0');
    for (const kg_hit& h : hits) {
      out << kmer[(size_t)h.idx] << '\t' << h.effect << '\t' << h.pval << '\n';
      if (do_2d && (h.flags & KG_HIT_2D)) {
        const string& b = bitstr[(size_t)h.idx];
        for (int j = 0; j < n; ++j) ob[j] = ((size_t)ov.idxs[j] < b.size()) ? b[ov.idxs[j]] : '0';
        out2 << kmer[(size_t)h.idx] << '\t' << ob << '\t' << h.effect << '\t' << h.pval << '\n';
      }
    }
    std::cerr << "[INFO] Processed " << lines << " kmers in " << tall.sec() << "s\n";
  }
  if (use_gz) gzclose(gz);
  std::cerr << "[INFO] Total kmers: " << lines << ", passed correlation filter: " << passed_total
            << "\n";
  if (do_2d) {
    out2.close();
    std::cerr << "[INFO] 2D hits written to " << path_2d(cfg.out) << "\n";
  }
  out.close();
  write_count(cfg.out, passed_total);
  std::cerr << "[TIME] End-to-end " << tall.sec() << "s\n";
}

}  // namespace

int main(int argc, char** argv) {
  Config cfg = parse_args(argc, argv);
  if (cfg.mode == "cols") run_cols(cfg);
  else run_rows(cfg);
  return 0;
}
