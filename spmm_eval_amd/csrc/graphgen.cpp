// Seeded synthetic graphs for the SpMM workloads: symmetric Chung-Lu
// power-law adjacency in CSR (no self-loops, no duplicate entries, column
// indices ascending in each row), vertex labels randomly permuted so hubs
// are spread over the row range as in real graphs.
//
// Vertex i (before permutation) has weight w_i = (i + i0)^(-1/(gamma-1));
// undirected edges are drawn with both endpoints ~ w (alias method),
// self-loops dropped, duplicates merged, both directions stored. i0 is
// solved so the expected maximum degree over the mean degree equals
// max_deg / (target_nnz / n). Pairs are drawn in fixed chunks, each from its
// own SplitMix64 stream, so the graph depends only on (n, target_nnz, gamma,
// max_deg, seed) and not on the thread count. If merging duplicates leaves
// fewer than target_nnz entries, more pairs are drawn (the first ones are
// unchanged) until the count is within 0.2% of the target or 6 rounds pass.
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
};

uint64_t mix(uint64_t a, uint64_t b) {
  SplitMix64 r(a ^ (b * 0xD1B54A32D192ED03ull));
  return r.next();
}

struct Alias {
  std::vector<double> prob;
  std::vector<int> alias;
  void build(const std::vector<double> &w) {
    const int n = static_cast<int>(w.size());
    prob.assign(n, 0.0);
    alias.assign(n, 0);
    double sum = 0;
    for (double x : w) sum += x;
    std::vector<double> p(n);
    std::vector<int> small, large;
    small.reserve(n);
    large.reserve(n);
    for (int i = 0; i < n; ++i) {
      p[i] = w[i] * n / sum;
      (p[i] < 1.0 ? small : large).push_back(i);
    }
    while (!small.empty() && !large.empty()) {
      int s = small.back(); small.pop_back();
      int l = large.back(); large.pop_back();
      prob[s] = p[s];
      alias[s] = l;
      p[l] = (p[l] + p[s]) - 1.0;
      (p[l] < 1.0 ? small : large).push_back(l);
    }
    for (int i : large) { prob[i] = 1.0; alias[i] = i; }
    for (int i : small) { prob[i] = 1.0; alias[i] = i; }
  }
  int draw(SplitMix64 &r) const {
    const int n = static_cast<int>(prob.size());
    uint64_t x = r.next();
    int i = static_cast<int>((x >> 32) * static_cast<uint64_t>(n) >> 32);
    double u = (x & 0xffffffffull) * (1.0 / 4294967296.0);
    return u < prob[i] ? i : alias[i];
  }
};

double max_over_mean(int n, double i0, double expo) {
  double sum = 0;
  for (int i = 0; i < n; ++i) sum += std::pow(i + i0, -expo);
  return std::pow(i0, -expo) * n / sum;
}

constexpr int64_t kChunk = 1 << 20;

}  // namespace

struct gen_graph_s {
  int n = 0;
  int64_t nnz = 0;
  int64_t pairs = 0;
  double i0 = 0;
  std::vector<int64_t> rowptr;
  std::vector<int> colind;
};

extern "C" {

typedef struct gen_graph_s gen_graph_t;

// Returns nullptr on bad arguments.
gen_graph_t *gen_chung_lu(int n, int64_t target_nnz, double gamma, double max_deg,
                          uint64_t seed) {
  if (n < 2 || target_nnz < 0 || gamma <= 1.0) return nullptr;
  auto *g = new gen_graph_s();
  g->n = n;
  const double expo = 1.0 / (gamma - 1.0);
  const double mean_deg = static_cast<double>(target_nnz) / n;
  double want = (max_deg > 0 && mean_deg > 0) ? max_deg / mean_deg : 0.0;
  // Solve i0: max_over_mean decreases in i0.
  double lo = 1e-3, hi = 1e9;
  if (want <= 1.0) {
    g->i0 = hi;
  } else if (max_over_mean(n, lo, expo) <= want) {
    g->i0 = lo;
  } else {
    for (int it = 0; it < 100; ++it) {
      double mid = std::sqrt(lo * hi);
      if (max_over_mean(n, mid, expo) > want) lo = mid; else hi = mid;
    }
    g->i0 = hi;
  }
  std::vector<double> w(n);
  for (int i = 0; i < n; ++i) w[i] = std::pow(i + g->i0, -expo);
  Alias alias;
  alias.build(w);

  // Random relabelling (Fisher-Yates).
  std::vector<int> perm(n);
  for (int i = 0; i < n; ++i) perm[i] = i;
  {
    SplitMix64 r(mix(seed, 0x5045524dull));
    for (int i = n - 1; i > 0; --i) {
      int j = static_cast<int>(r.next() % static_cast<uint64_t>(i + 1));
      std::swap(perm[i], perm[j]);
    }
  }

  int64_t pairs = target_nnz / 2;
  if (pairs < 1) pairs = 1;
  std::vector<int> U, V;
  for (int round = 0; round < 6; ++round) {
    const int64_t have = static_cast<int64_t>(U.size());
    U.resize(pairs);
    V.resize(pairs);
    const int64_t c0 = have / kChunk;  // chunks are regenerated from the first partial one
    const int64_t c1 = (pairs + kChunk - 1) / kChunk;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t c = c0; c < c1; ++c) {
      SplitMix64 r(mix(seed, static_cast<uint64_t>(c) + 1));
      const int64_t b = c * kChunk, e = std::min(pairs, b + kChunk);
      for (int64_t k = b; k < e; ++k) {
        int u = alias.draw(r), v = alias.draw(r);
        U[k] = perm[u];
        V[k] = perm[v];
      }
    }
    // Degrees with duplicates, both directions, self-loops dropped.
    std::vector<int64_t> cnt(n + 1, 0);
    for (int64_t k = 0; k < pairs; ++k) {
      if (U[k] == V[k]) continue;
      cnt[U[k]]++;
      cnt[V[k]]++;
    }
    std::vector<int64_t> ptr(n + 1, 0);
    for (int i = 0; i < n; ++i) ptr[i + 1] = ptr[i] + cnt[i];
    std::vector<int> col(ptr[n]);
    std::vector<int64_t> cur(ptr.begin(), ptr.end() - 1);
    for (int64_t k = 0; k < pairs; ++k) {
      int u = U[k], v = V[k];
      if (u == v) continue;
      col[cur[u]++] = v;
      col[cur[v]++] = u;
    }
    std::vector<int64_t> deg(n, 0);
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
      int *b = col.data() + ptr[i], *e = col.data() + ptr[i + 1];
      std::sort(b, e);
      deg[i] = std::unique(b, e) - b;
    }
    g->rowptr.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) g->rowptr[i + 1] = g->rowptr[i] + deg[i];
    g->nnz = g->rowptr[n];
    g->colind.resize(g->nnz);
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i)
      std::copy(col.begin() + ptr[i], col.begin() + ptr[i] + deg[i],
                g->colind.begin() + g->rowptr[i]);
    g->pairs = pairs;
    if (target_nnz == 0 || g->nnz >= target_nnz * 0.998) break;
    // Distinct entries per drawn pair so far; draw the shortfall at that rate.
    double rate = static_cast<double>(g->nnz) / pairs;
    if (rate <= 0) break;
    pairs += static_cast<int64_t>(std::ceil((target_nnz - g->nnz) / rate * 1.01));
  }
  return g;
}

int gen_graph_n(const gen_graph_t *g) { return g ? g->n : -1; }
int64_t gen_graph_nnz(const gen_graph_t *g) { return g ? g->nnz : -1; }
int64_t gen_graph_pairs(const gen_graph_t *g) { return g ? g->pairs : -1; }
double gen_graph_i0(const gen_graph_t *g) { return g ? g->i0 : -1; }

// rowptr must hold n+1 int32 (nnz < 2^31 required), colind nnz int32.
int gen_graph_copy(const gen_graph_t *g, int *rowptr, int *colind) {
  if (!g || !rowptr || !colind) return 1;
  if (g->nnz >= (int64_t(1) << 31)) return 2;
  for (int i = 0; i <= g->n; ++i) rowptr[i] = static_cast<int>(g->rowptr[i]);
  memcpy(colind, g->colind.data(), sizeof(int) * g->nnz);
  return 0;
}

void gen_graph_free(gen_graph_t *g) { delete g; }

}  // extern "C"
