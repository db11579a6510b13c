// Host-side CSR construction, validation, relabelling, file I/O and seeded
// synthetic generators (declared in include/acc_graph.h).
//
// Format and validation semantics follow the reference loader
// (scan-query/util/graph.cpp:48-143) and converter
// (scan-query/converter/EdgeListToBinsConverter.cpp:24-126).

#include "acc_graph.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>
#include <string>
#include <vector>

struct acc_hgraph {
  uint32_t n = 0;
  uint64_t m = 0;
  std::vector<uint32_t> off;
  std::vector<int32_t> adj;
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

// ---------------------------------------------------------------- RNG
// SplitMix64 (Steele, Lea, Flood 2014): the output mix is a bijection on
// 64-bit words, the state advances by the golden-ratio increment.
inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

struct SampleRng {
  uint64_t s;
  SampleRng(uint64_t seed, uint64_t sample) {
    s = mix64(seed * 0x9E3779B97F4A7C15ULL + 0x632BE59BD9B4E019ULL) ^
        mix64((sample + 1) * 0xD1B54A32D192ED03ULL);
  }
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ULL;
    return mix64(s);
  }
  // uniform in [0, 1) with 53 random bits
  double uniform() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); }
  // uniform integer in [0, n) by 64x64->128 multiply-high (Lemire 2019, no rejection;
  // the bias is below 2^-32 for n < 2^32 and the result is deterministic)
  uint32_t below(uint32_t n) {
    return static_cast<uint32_t>((static_cast<unsigned __int128>(next()) * n) >> 64);
  }
};

// ---------------------------------------------------------------- radix sort
// Parallel LSD radix sort of 64-bit keys whose significant bits are < key_bits.
void radix_sort_u64(std::vector<uint64_t> &keys, int key_bits) {
  const size_t N = keys.size();
  if (N < 2) return;
  std::vector<uint64_t> tmp(N);
  const int DIG = 11;
  const int B = 1 << DIG;
  int nthreads = omp_get_max_threads();
  std::vector<size_t> hist(static_cast<size_t>(nthreads) * B);
  uint64_t *src = keys.data();
  uint64_t *dst = tmp.data();
  int passes = 0;
  for (int shift = 0; shift < key_bits; shift += DIG) {
    std::fill(hist.begin(), hist.end(), 0);
#pragma omp parallel num_threads(nthreads)
    {
      int t = omp_get_thread_num();
      size_t lo = N * t / nthreads, hi = N * (t + 1) / nthreads;
      size_t *h = &hist[static_cast<size_t>(t) * B];
      for (size_t i = lo; i < hi; i++) h[(src[i] >> shift) & (B - 1)]++;
#pragma omp barrier
#pragma omp single
      {
        size_t run = 0;
        for (int d = 0; d < B; d++)
          for (int tt = 0; tt < nthreads; tt++) {
            size_t c = hist[static_cast<size_t>(tt) * B + d];
            hist[static_cast<size_t>(tt) * B + d] = run;
            run += c;
          }
      }
      for (size_t i = lo; i < hi; i++) dst[h[(src[i] >> shift) & (B - 1)]++] = src[i];
    }
    std::swap(src, dst);
    passes++;
  }
  if (passes & 1) keys.swap(tmp);
}

int bits_for(uint64_t x) {
  int b = 1;
  while (b < 64 && (x >> b) != 0) b++;
  return b;
}

// Directed keys src*n+dst (both directions), sorted and unique -> CSR.
int csr_from_directed_keys(uint32_t n, std::vector<uint64_t> &keys, acc_hgraph **out) {
  radix_sort_u64(keys, bits_for(static_cast<uint64_t>(n) * n));
  size_t w = 0;
  for (size_t i = 0; i < keys.size(); i++)
    if (i == 0 || keys[i] != keys[i - 1]) keys[w++] = keys[i];
  keys.resize(w);
  if (w >= (1ULL << 32)) return fail(ACC_EINVAL, "graph has 2^32 or more directed slots");
  auto *g = new (std::nothrow) acc_hgraph;
  if (!g) return fail(ACC_ENOMEM, "host allocation failed");
  g->n = n;
  g->m = w;
  g->off.assign(static_cast<size_t>(n) + 1, 0);
  g->adj.resize(w);
  for (size_t i = 0; i < w; i++) g->off[keys[i] / n + 1]++;
  for (uint32_t i = 0; i < n; i++) g->off[i + 1] += g->off[i];
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < w; i++) g->adj[i] = static_cast<int32_t>(keys[i] % n);
  *out = g;
  return ACC_OK;
}

void push_pair(std::vector<uint64_t> &keys, size_t slot, uint32_t n, uint32_t u, uint32_t v) {
  keys[2 * slot] = static_cast<uint64_t>(u) * n + v;
  keys[2 * slot + 1] = static_cast<uint64_t>(v) * n + u;
}

// Keys for self-loop samples are marked with an all-ones value and removed.
const uint64_t kDrop = ~0ULL;

int finish_from_samples(uint32_t n, std::vector<uint64_t> &keys, acc_hgraph **out) {
  size_t w = 0;
  for (size_t i = 0; i < keys.size(); i++)
    if (keys[i] != kDrop) keys[w++] = keys[i];
  keys.resize(w);
  return csr_from_directed_keys(n, keys, out);
}

}  // namespace

extern "C" {

const char *acc_graph_last_error(void) { return g_err.c_str(); }

uint32_t acc_hgraph_n(const acc_hgraph *g) { return g ? g->n : 0; }
uint64_t acc_hgraph_m(const acc_hgraph *g) { return g ? g->m : 0; }
const uint32_t *acc_hgraph_offsets(const acc_hgraph *g) { return g ? g->off.data() : nullptr; }
const int32_t *acc_hgraph_adj(const acc_hgraph *g) { return g ? g->adj.data() : nullptr; }
void acc_hgraph_free(acc_hgraph *g) { delete g; }

int acc_csr_validate(uint32_t n, uint64_t m, const uint32_t *off, const int32_t *adj) {
  if (!off || (m && !adj)) return fail(ACC_EINVAL, "null CSR pointer");
  if (off[0] != 0) return fail(ACC_EFORMAT, "node_off[0] != 0");
  if (off[n] != m) return fail(ACC_EFORMAT, "node_off[n] != m");
  for (uint32_t u = 0; u < n; u++)
    if (off[u + 1] < off[u]) return fail(ACC_EFORMAT, "node_off not monotone at " + std::to_string(u));
  // The error reported is the one of the smallest offending vertex.
  int bad = 0;
  long long bad_u = -1;
#pragma omp parallel for schedule(dynamic, 1024)
  for (long long u = 0; u < static_cast<long long>(n); u++) {
    for (uint64_t j = off[u]; j < off[u + 1]; j++) {
      int32_t v = adj[j];
      int code = 0;
      if (v < 0 || static_cast<uint32_t>(v) >= n) code = 1;
      else if (v == u) code = 2;
      else if (j > off[u] && adj[j] <= adj[j - 1]) code = 3;
      else if (!std::binary_search(adj + off[v], adj + off[v + 1], static_cast<int32_t>(u))) code = 4;
      if (code) {
#pragma omp critical
        {
          if (bad_u < 0 || u < bad_u) {
            bad_u = u;
            bad = code;
          }
        }
        break;
      }
    }
  }
  if (bad == 1) return fail(ACC_EFORMAT, "neighbor id out of range near vertex " + std::to_string(bad_u));
  if (bad == 2) return fail(ACC_EFORMAT, "Self loop near vertex " + std::to_string(bad_u));
  if (bad == 3) return fail(ACC_EFORMAT, "Edges not sorted in increasing id order near vertex " + std::to_string(bad_u));
  if (bad == 4) return fail(ACC_EFORMAT, "CSR not symmetric near vertex " + std::to_string(bad_u));
  return ACC_OK;
}

int acc_hgraph_from_edges(uint32_t n, uint64_t num_edges, const int32_t *src, const int32_t *dst,
                          acc_hgraph **out) {
  if (!out || n == 0 || (num_edges && (!src || !dst))) return fail(ACC_EINVAL, "bad argument");
  std::vector<uint64_t> keys;
  try {
    keys.resize(2 * num_edges);
  } catch (...) {
    return fail(ACC_ENOMEM, "host allocation failed");
  }
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (long long i = 0; i < static_cast<long long>(num_edges); i++) {
    int32_t s = src[i], d = dst[i];
    if (s < 0 || d < 0 || static_cast<uint32_t>(s) >= n || static_cast<uint32_t>(d) >= n) {
      bad = 1;
      keys[2 * i] = keys[2 * i + 1] = kDrop;
    } else if (s == d) {
      keys[2 * i] = keys[2 * i + 1] = kDrop;
    } else {
      push_pair(keys, i, n, s, d);
    }
  }
  if (bad) return fail(ACC_EINVAL, "edge endpoint out of range");
  return finish_from_samples(n, keys, out);
}

int acc_hgraph_from_csr(uint32_t n, uint64_t m, const uint32_t *node_off, const int32_t *edge_dst,
                        acc_hgraph **out) {
  if (!out) return fail(ACC_EINVAL, "null out");
  int rc = acc_csr_validate(n, m, node_off, edge_dst);
  if (rc) return rc;
  auto *g = new (std::nothrow) acc_hgraph;
  if (!g) return fail(ACC_ENOMEM, "host allocation failed");
  g->n = n;
  g->m = m;
  g->off.assign(node_off, node_off + n + 1);
  g->adj.assign(edge_dst, edge_dst + m);
  *out = g;
  return ACC_OK;
}

int acc_hgraph_relabel_degree(acc_hgraph *g, int32_t *old_of_new_out) {
  if (!g) return fail(ACC_EINVAL, "null graph");
  const uint32_t n = g->n;
  // Stable counting sort by degree, descending.
  uint32_t maxd = 0;
  for (uint32_t u = 0; u < n; u++) maxd = std::max(maxd, g->off[u + 1] - g->off[u]);
  std::vector<uint32_t> bucket(static_cast<size_t>(maxd) + 2, 0);
  for (uint32_t u = 0; u < n; u++) bucket[maxd - (g->off[u + 1] - g->off[u]) + 1]++;
  for (size_t i = 1; i < bucket.size(); i++) bucket[i] += bucket[i - 1];
  std::vector<int32_t> old_of_new(n), new_of_old(n);
  for (uint32_t u = 0; u < n; u++) {
    uint32_t pos = bucket[maxd - (g->off[u + 1] - g->off[u])]++;
    old_of_new[pos] = static_cast<int32_t>(u);
    new_of_old[u] = static_cast<int32_t>(pos);
  }
  std::vector<uint32_t> noff(static_cast<size_t>(n) + 1, 0);
  for (uint32_t x = 0; x < n; x++) {
    uint32_t o = old_of_new[x];
    noff[x + 1] = noff[x] + (g->off[o + 1] - g->off[o]);
  }
  std::vector<int32_t> nadj(g->m);
#pragma omp parallel for schedule(dynamic, 256)
  for (long long x = 0; x < static_cast<long long>(n); x++) {
    uint32_t o = old_of_new[x];
    uint64_t w = noff[x];
    for (uint64_t j = g->off[o]; j < g->off[o + 1]; j++) nadj[w++] = new_of_old[g->adj[j]];
    std::sort(nadj.begin() + noff[x], nadj.begin() + noff[x + 1]);
  }
  g->off.swap(noff);
  g->adj.swap(nadj);
  if (old_of_new_out) std::memcpy(old_of_new_out, old_of_new.data(), sizeof(int32_t) * n);
  return ACC_OK;
}

int acc_gen_er(uint32_t n, uint64_t num_samples, uint64_t seed, acc_hgraph **out) {
  if (!out || n < 2) return fail(ACC_EINVAL, "ER needs n >= 2");
  std::vector<uint64_t> keys(2 * num_samples);
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < static_cast<long long>(num_samples); i++) {
    SampleRng r(seed, i);
    uint32_t u, v;
    do {
      u = r.below(n);
      v = r.below(n);
    } while (u == v);
    push_pair(keys, i, n, u, v);
  }
  return finish_from_samples(n, keys, out);
}

int acc_gen_rmat(uint32_t n, int scale, uint64_t num_samples, double a, double b, double c,
                 uint64_t seed, acc_hgraph **out) {
  if (!out || scale < 1 || scale > 31 || n == 0 || static_cast<uint64_t>(n) > (1ULL << scale))
    return fail(ACC_EINVAL, "R-MAT needs 0 < n <= 2^scale, scale in [1,31]");
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return fail(ACC_EINVAL, "bad R-MAT probabilities");
  const double ab = a + b, abc = a + b + c;
  std::vector<uint64_t> keys(2 * num_samples);
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < static_cast<long long>(num_samples); i++) {
    SampleRng r(seed, i);
    uint32_t u, v;
    do {
      u = 0;
      v = 0;
      for (int lvl = scale - 1; lvl >= 0; lvl--) {
        double x = r.uniform();
        // quadrant: a = (0,0), b = (0,1), c = (1,0), d = (1,1) as (row bit, column bit)
        uint32_t row = x >= ab ? 1u : 0u;
        uint32_t col = (x >= a && x < ab) || x >= abc ? 1u : 0u;
        u |= row << lvl;
        v |= col << lvl;
      }
    } while (u >= n || v >= n);
    if (u == v) {
      keys[2 * i] = keys[2 * i + 1] = kDrop;
    } else {
      push_pair(keys, i, n, u, v);
    }
  }
  return finish_from_samples(n, keys, out);
}

int acc_gen_chung_lu(uint32_t n, double gamma, double avg_deg, double max_deg, uint64_t seed,
                     acc_hgraph **out) {
  if (!out || n < 2 || gamma <= 1.0 || avg_deg <= 0 || max_deg <= 0)
    return fail(ACC_EINVAL, "bad Chung-Lu parameters");
  const double ex = -1.0 / (gamma - 1.0);
  std::vector<double> w(n);
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < static_cast<long long>(n); i++) w[i] = std::pow(static_cast<double>(i + 1), ex);
  // Find the scale s with mean(min(s*w_i, max_deg)) == avg_deg by bisection.
  auto mean_at = [&](double s) {
    double acc = 0;
#pragma omp parallel for reduction(+ : acc) schedule(static)
    for (long long i = 0; i < static_cast<long long>(n); i++) acc += std::min(s * w[i], max_deg);
    return acc / n;
  };
  double lo = 0, hi = 1;
  while (mean_at(hi) < avg_deg && hi < 1e30) hi *= 2;
  for (int it = 0; it < 200; it++) {
    double mid = 0.5 * (lo + hi);
    (mean_at(mid) < avg_deg ? lo : hi) = mid;
  }
  std::vector<double> cdf(static_cast<size_t>(n) + 1, 0.0);
  for (uint32_t i = 0; i < n; i++) cdf[i + 1] = cdf[i] + std::min(hi * w[i], max_deg);
  const double total = cdf[n];
  const uint64_t samples = static_cast<uint64_t>(std::llround(static_cast<double>(n) * avg_deg / 2.0));
  std::vector<uint64_t> keys(2 * samples);
  auto draw = [&](SampleRng &r) {
    double x = r.uniform() * total;
    size_t k = std::upper_bound(cdf.begin() + 1, cdf.end(), x) - (cdf.begin() + 1);
    return static_cast<uint32_t>(std::min<size_t>(k, n - 1));
  };
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < static_cast<long long>(samples); i++) {
    SampleRng r(seed, i);
    uint32_t u = draw(r), v = draw(r);
    if (u == v) {
      keys[2 * i] = keys[2 * i + 1] = kDrop;
    } else {
      push_pair(keys, i, n, u, v);
    }
  }
  return finish_from_samples(n, keys, out);
}

int acc_hgraph_write_bins(const acc_hgraph *g, const char *dir) {
  if (!g || !dir) return fail(ACC_EINVAL, "bad argument");
  if (g->m >= (1ULL << 32)) return fail(ACC_EINVAL, "m does not fit the uint32 header");
  std::string d(dir);
  FILE *f = std::fopen((d + "/b_degree.bin").c_str(), "wb");
  if (!f) return fail(ACC_EIO, "cannot write " + d + "/b_degree.bin");
  int32_t int_size = 4;
  uint32_t n = g->n, m = static_cast<uint32_t>(g->m);
  std::vector<int32_t> deg(n);
  for (uint32_t u = 0; u < n; u++) deg[u] = static_cast<int32_t>(g->off[u + 1] - g->off[u]);
  bool ok = std::fwrite(&int_size, 4, 1, f) == 1 && std::fwrite(&n, 4, 1, f) == 1 &&
            std::fwrite(&m, 4, 1, f) == 1 && (n == 0 || std::fwrite(deg.data(), 4, n, f) == n);
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(ACC_EIO, "short write to b_degree.bin");
  f = std::fopen((d + "/b_adj.bin").c_str(), "wb");
  if (!f) return fail(ACC_EIO, "cannot write " + d + "/b_adj.bin");
  ok = g->m == 0 || std::fwrite(g->adj.data(), 4, g->m, f) == g->m;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(ACC_EIO, "short write to b_adj.bin");
  return ACC_OK;
}

int acc_hgraph_read_bins(const char *dir, acc_hgraph **out) {
  if (!dir || !out) return fail(ACC_EINVAL, "bad argument");
  std::string d(dir);
  FILE *f = std::fopen((d + "/b_degree.bin").c_str(), "rb");
  if (!f) return fail(ACC_EIO, "cannot open " + d + "/b_degree.bin");
  int32_t int_size = 0;
  uint32_t n = 0, m = 0;
  bool ok = std::fread(&int_size, 4, 1, f) == 1 && std::fread(&n, 4, 1, f) == 1 &&
            std::fread(&m, 4, 1, f) == 1;
  std::vector<int32_t> deg;
  if (ok) {
    deg.resize(n);
    ok = n == 0 || std::fread(deg.data(), 4, n, f) == n;
  }
  std::fclose(f);
  if (!ok || int_size != 4) return fail(ACC_EFORMAT, "malformed b_degree.bin");
  auto *g = new (std::nothrow) acc_hgraph;
  if (!g) return fail(ACC_ENOMEM, "host allocation failed");
  g->n = n;
  g->m = m;
  g->off.assign(static_cast<size_t>(n) + 1, 0);
  for (uint32_t u = 0; u < n; u++) {
    if (deg[u] < 0) {
      delete g;
      return fail(ACC_EFORMAT, "negative degree in b_degree.bin");
    }
    g->off[u + 1] = g->off[u] + static_cast<uint32_t>(deg[u]);
  }
  if (g->off[n] != m) {
    delete g;
    return fail(ACC_EFORMAT, "degrees in b_degree.bin do not sum to m");
  }
  g->adj.resize(m);
  f = std::fopen((d + "/b_adj.bin").c_str(), "rb");
  if (!f) {
    delete g;
    return fail(ACC_EIO, "cannot open " + d + "/b_adj.bin");
  }
  ok = m == 0 || std::fread(g->adj.data(), 4, m, f) == m;
  std::fclose(f);
  if (!ok) {
    delete g;
    return fail(ACC_EFORMAT, "b_adj.bin shorter than m entries");
  }
  int rc = acc_csr_validate(n, m, g->off.data(), g->adj.data());
  if (rc) {
    delete g;
    return rc;
  }
  *out = g;
  return ACC_OK;
}

int acc_convert_edge_list(const char *txt_path, const char *out_dir) {
  if (!txt_path || !out_dir) return fail(ACC_EINVAL, "bad argument");
  std::ifstream ifs(txt_path);
  if (!ifs) return fail(ACC_EIO, std::string("cannot open ") + txt_path);
  std::vector<int32_t> s, t;
  long long max_ele = -1;
  std::string line;
  while (std::getline(ifs, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ss(line);
    long long a, b;
    if (!(ss >> a >> b)) continue;
    if (a >= b) continue;  // the converter keeps only first < second
    if (a < 0 || b >= INT32_MAX) return fail(ACC_EFORMAT, "vertex id out of int32 range");
    max_ele = std::max(max_ele, b);
    s.push_back(static_cast<int32_t>(a));
    t.push_back(static_cast<int32_t>(b));
  }
  if (max_ele < 0) return fail(ACC_EFORMAT, "no edges with first < second");
  acc_hgraph *g = nullptr;
  int rc = acc_hgraph_from_edges(static_cast<uint32_t>(max_ele + 1), s.size(), s.data(), t.data(), &g);
  if (rc) return rc;
  rc = acc_hgraph_write_bins(g, out_dir);
  acc_hgraph_free(g);
  return rc;
}

int acc_write_scan_result(const char *path, uint32_t n, const int32_t *label, const int32_t *cluster_id,
                          const int32_t *old_id, int32_t *pairs, uint64_t npairs) {
  if (!path || (n && (!label || !cluster_id)) || (npairs && !pairs)) return fail(ACC_EINVAL, "bad argument");
  std::vector<std::pair<int32_t, int32_t>> cores;
  for (uint32_t u = 0; u < n; u++)
    if (label[u] == 1) cores.emplace_back(old_id ? old_id[u] : static_cast<int32_t>(u), cluster_id[u]);
  std::sort(cores.begin(), cores.end());
  auto *pp = reinterpret_cast<std::pair<int32_t, int32_t> *>(pairs);
  static_assert(sizeof(std::pair<int32_t, int32_t>) == 8, "pair layout");
  std::sort(pp, pp + npairs);
  FILE *f = std::fopen(path, "w");
  if (!f) return fail(ACC_EIO, std::string("cannot write ") + path);
  std::string buf;
  buf.reserve(1 << 20);
  buf += "c/n vertex_id cluster_id\n";
  char line[64];
  bool ok = true;
  auto flush = [&]() {
    if (!buf.empty() && std::fwrite(buf.data(), 1, buf.size(), f) != buf.size()) ok = false;
    buf.clear();
  };
  for (auto &c : cores) {
    std::snprintf(line, sizeof(line), "c %d %d\n", c.first, c.second);
    buf += line;
    if (buf.size() > (1 << 20)) flush();
  }
  for (uint64_t i = 0; i < npairs; i++) {
    if (i > 0 && pp[i] == pp[i - 1]) continue;
    std::snprintf(line, sizeof(line), "n %d %d\n", pp[i].second, pp[i].first);
    buf += line;
    if (buf.size() > (1 << 20)) flush();
  }
  flush();
  ok = (std::fclose(f) == 0) && ok;
  return ok ? ACC_OK : fail(ACC_EIO, std::string("short write to ") + path);
}

int acc_scan_result_path(const char *dir, double eps, int mu, char *buf, int cap) {
  if (!dir || !buf || cap <= 0) return fail(ACC_EINVAL, "bad argument");
  std::string e = std::to_string(eps);
  e.erase(e.find_last_not_of("0u") + 1);
  std::string p = std::string(dir) + "/scanxp-result-" + e + "-" + std::to_string(mu) + ".txt";
  if (static_cast<int>(p.size()) >= cap) return fail(ACC_EINVAL, "buffer too small");
  std::memcpy(buf, p.c_str(), p.size() + 1);
  return ACC_OK;
}

int acc_hgraph_reverse_slots(const acc_hgraph *g, uint32_t *rev) {
  if (!g || !rev) return fail(ACC_EINVAL, "bad argument");
#pragma omp parallel for schedule(dynamic, 1024)
  for (long long u = 0; u < static_cast<long long>(g->n); u++) {
    for (uint64_t j = g->off[u]; j < g->off[u + 1]; j++) {
      int32_t v = g->adj[j];
      const int32_t *b = g->adj.data() + g->off[v], *e = g->adj.data() + g->off[v + 1];
      rev[j] = static_cast<uint32_t>(std::lower_bound(b, e, static_cast<int32_t>(u)) - g->adj.data());
    }
  }
  return ACC_OK;
}

int acc_csr_verify_counts(uint32_t n, uint64_t m, const uint32_t *off, const int32_t *adj, const int32_t *cnt,
                          int32_t base, int nthreads, uint64_t *mismatches, uint64_t *first_bad) {
  if (!off || (m && (!adj || !cnt)) || !mismatches || !first_bad) return fail(ACC_EINVAL, "bad argument");
  if (off[0] != 0 || off[n] != m) return fail(ACC_EINVAL, "node_off must run from 0 to m");
  const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
  const uint64_t words = (static_cast<uint64_t>(n) + 63) / 64;
  uint64_t bad = 0, first = m;
#pragma omp parallel num_threads(nt) reduction(+ : bad) reduction(min : first)
  {
    // this thread's bit vector of N(u): set for u, probe every N(v) with v > u, clear again
    std::vector<uint64_t> bits(words ? words : 1, 0);
#pragma omp for schedule(dynamic, 64)
    for (long long u = 0; u < static_cast<long long>(n); u++) {
      const uint64_t s = off[u], e = off[u + 1];
      for (uint64_t j = s; j < e; j++) bits[static_cast<uint32_t>(adj[j]) >> 6] |= 1ULL << (adj[j] & 63);
      for (uint64_t j = s; j < e; j++) {
        const int32_t v = adj[j];
        if (v <= u || static_cast<uint32_t>(v) >= n) continue;
        int32_t c = base;
        for (uint64_t k = off[v]; k < off[v + 1]; k++)
          c += static_cast<int32_t>((bits[static_cast<uint32_t>(adj[k]) >> 6] >> (adj[k] & 63)) & 1);
        if (cnt[j] != c) {
          bad++;
          first = std::min<uint64_t>(first, j);
        }
      }
      for (uint64_t j = s; j < e; j++) bits[static_cast<uint32_t>(adj[j]) >> 6] = 0;
    }
  }
  *mismatches = bad;
  *first_bad = first;
  return ACC_OK;
}

}  // extern "C"
