// kgwas_2d -- drop-in for SuperGWAS's all-pairs epistasis binary, computing on
// one MI355X through libkgwas_hip (kg_scan_pairs, include/kgwas_hip.h).
//
//   kgwas_2d samples.txt phenotype.txt U2D.bin INDEX_DIR out_prefix
//            [log10_thresh] [--emit-threads N] [--out-chunk N]
//            [--task-id N] [--num-tasks N]
//
// Command line, environment (BLAS_THREADS, EMIT_THREADS), the d4 index it
// reads, its outputs ("{out_prefix}.tsv", or ".part_<task>.tsv" with several
// tasks, and "{out_prefix}.kmers" from task 0) and error behaviour
// ("[ERROR] msg" + exit 1) follow src/kgwas_2d.cpp:43-73, 75-203, 258-481.
// Additional environment: KG_DEVICE (HIP device, default 0).
//
// Numerics: every reported pair's effect and p come from FP64 values computed
// on the GPU and the reference's p-value expression on the host, where the
// reference forms them in FP32 (src/kgwas_2d.cpp:219-251, 372-382); digits
// beyond FP32's can therefore differ.
#include <algorithm>
#include <cstring>
#include <iomanip>
#include <sstream>

#include "kgwas_hip.h"
#include "kgwas_io.hpp"

using namespace kgio;

namespace {

struct Config {
  string samples, pheno, ubin, index_dir, out_prefix;
  double log10_thresh = 6.0;
  int blas_threads = 112, emit_threads = 112, out_chunk = 1000;
  int task_id = 0, num_tasks = 1;
};

int parse_int(const string& s, const string& what) {
  try {
    return std::stoi(s);
  } catch (...) {
    die("invalid " + what + ": " + s);
  }
}

Config parse_args(int argc, char** argv) {  // src/kgwas_2d.cpp:43-73
  if (argc < 6) {
    std::cerr << "Usage: " << argv[0]
              << " samples.txt phenotype.txt U2D.bin INDEX_DIR out_prefix"
              << " [log10_thresh] [--emit-threads N] [--out-chunk N]\n"
              << "\nOutputs:\n"
              << "  out_prefix.tsv   — idx_a, idx_b, effect, pvalue\n"
              << "  out_prefix.kmers — one k-mer sequence per line (line# = index)\n";
    std::exit(1);
  }
  Config c;
  c.samples = argv[1];
  c.pheno = argv[2];
  c.ubin = argv[3];
  c.index_dir = argv[4];
  c.out_prefix = argv[5];
  const bool thr = argc > 6 && argv[6][0] != '-';
  if (thr) {
    try {
      c.log10_thresh = std::stod(argv[6]);
    } catch (...) {
      die(string("invalid log10_thresh: ") + argv[6]);
    }
  }
  for (int i = thr ? 7 : 6; i < argc; ++i) {
    const string a = argv[i];
    if (a == "--emit-threads" && i + 1 < argc) c.emit_threads = parse_int(argv[++i], a);
    else if (a == "--out-chunk" && i + 1 < argc) c.out_chunk = parse_int(argv[++i], a);
    else if (a == "--task-id" && i + 1 < argc) c.task_id = parse_int(argv[++i], a);
    else if (a == "--num-tasks" && i + 1 < argc) c.num_tasks = parse_int(argv[++i], a);
  }
  c.blas_threads = env_int("BLAS_THREADS", c.blas_threads);
  c.emit_threads = env_int("EMIT_THREADS", c.emit_threads);
  if (c.num_tasks < 1) die("--num-tasks must be >= 1");
  if (c.task_id < 0) die("--task-id must be >= 0");
  return c;
}

// ---- sharded d4 index (src/map_utils.hpp:319-337, 382-479) -----------------
//   index.meta           text lines K=, M=, N=
//   shard_%03u.umap      "UMAPd4\0\0", uint32 K, uint64 N, then N records of
//                        uint128 encoded k-mer, float64 effect, float64 pval
//   shard_%03u.bits      uint64 N, then per record uint32 length + that many
//                        '0'/'1' bytes (optional when no record has one)
typedef unsigned __int128 u128;

struct U128Hash {
  size_t operator()(u128 k) const noexcept {
    const uint64_t lo = (uint64_t)k, hi = (uint64_t)(k >> 64);
    return std::hash<uint64_t>()(lo ^ (hi * 0x9E3779B97F4A7C15ull));
  }
};

void read_meta(const string& dir, uint32_t& K, uint32_t& M) {
  std::ifstream in(dir + "/index.meta");
  if (!in) die("cannot read index.meta in " + dir);
  string line;
  K = 0;
  M = 0;
  while (std::getline(in, line)) {
    try {
      if (line.compare(0, 2, "K=") == 0) K = (uint32_t)std::stoul(line.substr(2));
      else if (line.compare(0, 2, "M=") == 0) M = (uint32_t)std::stoul(line.substr(2));
    } catch (...) {
      die("invalid index.meta line: " + line);
    }
  }
  if (K == 0 || M == 0)
    die("invalid index.meta (K=" + std::to_string(K) + " M=" + std::to_string(M) + ")");
}

vector<char> slurp(const string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) die("cannot open " + path);
  in.seekg(0, std::ios::end);
  const size_t sz = (size_t)in.tellg();
  in.seekg(0, std::ios::beg);
  vector<char> buf(sz);
  if (sz && !in.read(buf.data(), (std::streamsize)sz)) die("short read: " + path);
  return buf;
}

// The k-mers of shard `id` that carry a bitstring.  A key repeated inside a
// shard is one entry whose bitstring is the last non-empty one read (the
// reference fills a hash map in file order, src/map_utils.hpp:427-475).
void load_shard(const string& dir, uint32_t id, uint32_t K, vector<std::pair<u128, string>>& out) {
  char name[32];
  std::snprintf(name, sizeof(name), "shard_%03u", id);
  const string up = dir + "/" + name + ".umap", bp = dir + "/" + name + ".bits";
  const vector<char> ub = slurp(up);
  if (ub.size() < 20) die("file too small: " + up);
  if (std::memcmp(ub.data(), "UMAPd4\0\0", 8) != 0) die("bad magic in " + up);
  uint32_t k;
  uint64_t N;
  std::memcpy(&k, ub.data() + 8, 4);
  std::memcpy(&N, ub.data() + 12, 8);
  if (k != K)
    die("K mismatch in " + up + ": got " + std::to_string(k) + " expected " + std::to_string(K));
  if (N > (ub.size() - 20) / 32) die("short read: " + up);
  const vector<char> bb = slurp(bp);
  if (bb.size() < 8) die("short read: " + bp);
  uint64_t bN;
  std::memcpy(&bN, bb.data(), 8);
  if (bN != N) die(".bits N=" + std::to_string(bN) + " != .umap N=" + std::to_string(N));
  std::unordered_map<u128, string, U128Hash> seen;
  vector<u128> order;
  size_t off = 8;
  for (uint64_t i = 0; i < N; ++i) {
    u128 enc;
    std::memcpy(&enc, ub.data() + 20 + 32 * i, 16);
    if (off + 4 > bb.size()) die("short read (bitstring len@" + std::to_string(i) + "): " + bp);
    uint32_t blen;
    std::memcpy(&blen, bb.data() + off, 4);
    off += 4;
    if (blen > bb.size() - off) die("short read (bitstring@" + std::to_string(i) + "): " + bp);
    auto it = seen.find(enc);
    if (it == seen.end()) {
      it = seen.emplace(enc, string()).first;
      order.push_back(enc);
    }
    if (blen > 0) it->second.assign(bb.data() + off, blen);
    off += blen;
  }
  for (u128 e : order) {
    auto it = seen.find(e);
    if (!it->second.empty()) out.emplace_back(e, std::move(it->second));
  }
}

bool is_dir(const string& p) {
  struct stat st {};
  return ::stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

string decode_kmer(u128 enc, uint32_t K) {  // src/map_utils.hpp:54-60
  static const char kBase[4] = {'A', 'C', 'G', 'T'};
  string s(K, 'A');
  for (uint32_t i = K; i > 0; --i) {
    s[i - 1] = kBase[(unsigned)(enc & 3)];
    enc >>= 2;
  }
  return s;
}

}  // namespace

int main(int argc, char** argv) {
  const Config cfg = parse_args(argc, argv);
  Clock tall;
  const Overlap ov = align_samples(cfg.samples, cfg.pheno);
  const int n = (int)ov.y.size();
  const UBin u = read_ubin(cfg.ubin, n);

  // index k-mers with bitstrings, sorted by encoded k-mer (src/kgwas_2d.cpp:85-150)
  Clock tload;
  uint32_t K = 0, M = 0;
  if (!is_dir(cfg.index_dir)) die("Index path is not a directory: " + cfg.index_dir);
  read_meta(cfg.index_dir, K, M);
  vector<std::pair<u128, string>> entries;
  for (uint32_t s = 0; s < M; ++s) load_shard(cfg.index_dir, s, K, entries);
  std::cerr << "[INFO] " << entries.size() << " k-mers with bitstrings (" << M
            << " shards, K=" << K << ", " << tload.sec() << " s)\n";
  if (entries.empty()) die("No k-mers with bitstrings in index");
  std::sort(entries.begin(), entries.end(),
            [](const std::pair<u128, string>& a, const std::pair<u128, string>& b) {
              return a.first < b.first;
            });

  // rows over the phenotyped samples in overlap order: a bitstring of length n
  // is read directly, a longer one at the accession positions of the overlap,
  // any other is dropped (src/kgwas_2d.cpp:159-190)
  const int maxidx = ov.idxs.empty() ? -1 : *std::max_element(ov.idxs.begin(), ov.idxs.end());
  const int64_t rw = ((n + 63) / 64 + 7) / 8 * 8;
  vector<uint64_t> rows;
  vector<string> kmers;
  rows.reserve(entries.size() * (size_t)rw);
  for (const auto& e : entries) {
    const string& b = e.second;
    const int B = (int)b.size();
    const bool direct = B == n, mapped = !direct && maxidx >= 0 && B > maxidx;
    if (!direct && !mapped) continue;
    const size_t base = rows.size();
    rows.resize(base + (size_t)rw, 0ull);
    for (int j = 0; j < n; ++j)
      if (b[direct ? j : ov.idxs[j]] == '1') rows[base + (j >> 6)] |= 1ull << (j & 63);
    kmers.push_back(decode_kmer(e.first, K));
  }
  const int64_t N = (int64_t)kmers.size();
  if (N <= 0) die("No valid k-mers after bitstring parsing");
  std::cerr << "[INFO] Built X from index (N=" << N << ", n=" << n << ")\n";
  {
    double yy = 0.0, uu = 0.0;  // src/kgwas_2d.cpp:241-243
    for (double v : ov.y) yy += v * v;
    for (int k = 0; k < u.r; ++k) {
      double s = 0.0;
      for (int j = 0; j < n; ++j) s += u.U[(size_t)k * n + j] * ov.y[j];
      uu += s * s;
    }
    if (!(yy - uu > 1e-12)) warn("yyM ~ 0");
  }
  if (N < 2) die("emit_pairs_self: need at least 2 columns");

  // this task's rows i (src/kgwas_2d.cpp:283-291)
  const int64_t per_task = (N + cfg.num_tasks - 1) / cfg.num_tasks;
  const int64_t row_start = (int64_t)cfg.task_id * per_task;
  const int64_t row_end = std::min(row_start + per_task, N);
  if (row_start >= N) {
    std::cerr << "[INFO] task " << cfg.task_id << ": no rows to process\n";
    std::cerr << "[TIME] total " << tall.sec() << " s\n";
    return 0;
  }

  // the pair scan on the GPU: the model's accessions are the phenotyped samples
  Clock tscan;
  kg_opts o{};
  const char* dv = std::getenv("KG_DEVICE");
  o.device = dv ? std::atoi(dv) : 0;
  o.precision = KG_PREC_MIXED;
  kg_ctx* ctx = nullptr;
  int rc = kg_create(&ctx, &o);
  if (rc != KG_OK) die(string("GPU init failed: ") + (ctx ? kg_last_error(ctx) : "kg_create"));
  vector<int32_t> idn(n);
  for (int j = 0; j < n; ++j) idn[j] = j;
  rc = kg_set_model(ctx, n, idn.data(), n, u.U.empty() ? nullptr : u.U.data(), u.r, ov.y.data(),
                    1);
  if (rc != KG_OK) die(string("model set-up failed: ") + kg_last_error(ctx));
  void* d_rows = nullptr;
  const size_t nbytes = sizeof(uint64_t) * rows.size();
  if (kg_dev_alloc(ctx, nbytes, &d_rows) != KG_OK ||
      kg_memcpy(ctx, d_rows, rows.data(), nbytes, 0) != KG_OK)
    die(string("device rows: ") + kg_last_error(ctx));
  const kg_scan_params prm{0.0, cfg.log10_thresh, 0.0};
  vector<kg_hit> pairs(1024);
  for (;;) {
    int64_t nout = 0;
    rc = kg_scan_pairs(ctx, static_cast<const uint64_t*>(d_rows), N, rw, row_start, row_end, &prm,
                       pairs.data(), (int64_t)pairs.size(), &nout);
    if (rc == KG_ERR_CAPACITY) {
      pairs.resize((size_t)nout);
      continue;
    }
    if (rc != KG_OK) die(string("pair scan failed: ") + kg_last_error(ctx));
    pairs.resize((size_t)nout);
    break;
  }
  kg_dev_free(ctx, d_rows);
  kg_destroy(ctx);
  std::cerr << "[INFO] found " << pairs.size() << " passing pairs in " << tscan.sec()
            << " s (rows [" << row_start << "," << row_end << ") of N=" << N << ")\n";

  // outputs (src/kgwas_2d.cpp:427-472)
  {
    string path = cfg.out_prefix;
    if (cfg.num_tasks > 1) path += ".part_" + std::to_string(cfg.task_id);
    path += ".tsv";
    std::ofstream out(path, std::ios::binary);
    if (!out) die("open out: " + path);
    std::ostringstream oss;
    oss.setf(std::ios::fmtflags(0), std::ios::floatfield);
    for (const kg_hit& h : pairs)
      oss << (int64_t)h.shard << '\t' << h.idx << '\t' << std::setprecision(9) << std::defaultfloat
          << h.effect << '\t' << std::setprecision(9) << std::scientific << h.pval << '\n';
    const string s = oss.str();
    out.write(s.data(), (std::streamsize)s.size());
    out.close();
    if (!out) die("write out failed");
    std::cerr << "[INFO] wrote " << pairs.size() << " pairs to " << path << "\n";
  }
  if (cfg.task_id == 0) {
    const string path = cfg.out_prefix + ".kmers";
    std::ofstream out(path);
    if (!out) die("open out: " + path);
    for (const string& k : kmers) out << k << '\n';
    out.close();
    std::cerr << "[INFO] wrote " << kmers.size() << " k-mers to " << path << "\n";
  }
  std::cerr << "[TIME] total " << tall.sec() << " s\n";
  return 0;
}
