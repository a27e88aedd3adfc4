// Seeded synthetic Ising instances built on the host, in parallel, straight into symmetric CSR: the
// inputs of the large workloads (Erdos-Renyi and Barabasi-Albert graphs with up to 2^25 nodes and
// ~7e8 stored couplings), which the vectorised numpy generators take tens of seconds to build.
//
// Every random quantity is a counter-based hash of (seed, salt, position), so the graph does not
// depend on the number of threads or on how rows are split among them:
//   * Erdos-Renyi G(n, p): row i's partners j > i by geometric skipping, draw k of row i from
//     hash(seed, i, k) (Batagelj & Brandes 2005, algorithm 1);
//   * Barabasi-Albert, m edges per new node by the edge-list copying model (Batagelj & Brandes 2005,
//     algorithm 5): slot 2k holds the source k / m of edge k, slot 2k + 1 copies slot r_k, r_k uniform
//     in [0, 2k]. Each odd slot is resolved independently by following r until it lands on an even
//     slot (expected two steps), so all edges are drawn in parallel without the slot array;
//     self-loops are dropped and repeated edges merged;
//   * couplings per undirected edge {i, j}: -1 ("neg") or +-1 from hash(seed, i, j) ("pm1").
// Rows are sorted by column and hold no duplicates; nodes left without couplings are dropped and the
// others renumbered in order (as the instance loader drops isolated nodes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <new>
#include <thread>
#include <vector>

#include "sa_common.hpp"

struct sa_graph {
  int64_t n = 0;
  std::vector<int64_t> indptr;
  std::vector<int32_t> indices;
  std::vector<double> vals;
};

namespace sa {
namespace {

constexpr uint64_t kErSalt = 0xE5D05E4E11ull;
constexpr uint64_t kBaSalt = 0xBA2BA511ull;
constexpr uint64_t kSignSalt = 0x5167ull;

// Run f(lo, hi) over [0, total) in chunks handed out dynamically to `threads` workers.
template <class F>
void parallel_chunks(int64_t total, int64_t chunk, int threads, F f) {
  std::atomic<int64_t> next{0};
  auto worker = [&]() {
    for (;;) {
      const int64_t lo = next.fetch_add(chunk);
      if (lo >= total) return;
      f(lo, std::min(total, lo + chunk));
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; ++t) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
}

double coupling(int32_t kind, uint64_t seed, int64_t i, int64_t j) {
  if (kind == 0) return -1.0;
  const int64_t a = i < j ? i : j, b = i < j ? j : i;
  return (mix3(seed ^ kSignSalt, (uint64_t)a, (uint64_t)b) & 1ull) ? 1.0 : -1.0;
}

// Row r's upper partners j > r of G(n, p), in increasing order, handed to emit(j).
template <class E>
void er_row(int64_t n, double p, double lq, uint64_t seed, int64_t r, E emit) {
  if (p <= 0.0) return;
  int64_t j = r;
  for (uint64_t k = 0;; ++k) {
    int64_t skip = 0;
    if (p < 1.0) {
      const double u = u53(mix3(seed ^ kErSalt, (uint64_t)r, k));
      const double g = std::floor(std::log1p(-u) / lq);
      if (!(g < (double)n)) return;
      skip = (int64_t)g;
    }
    j += 1 + skip;
    if (j >= n) return;
    emit(j);
  }
}

// The node slot 2k + 1 of the Barabasi-Albert slot array resolves to.
int64_t ba_target(int64_t k, int32_t m, uint64_t seed) {
  for (;;) {
    const uint64_t span = 2ull * (uint64_t)k + 1ull;  // r uniform in [0, 2k]
    uint64_t r = (uint64_t)(u53(mix3(seed ^ kBaSalt, (uint64_t)k, 0)) * (double)span);
    if (r >= span) r = span - 1;
    if ((r & 1ull) == 0) return (int64_t)(r >> 1) / m;  // slot 2k' holds the source of edge k'
    k = (int64_t)(r >> 1);                              // slot 2k' + 1: follow its copy
  }
}

// Symmetric CSR from per-row counts of the undirected edges each row owns (owner rows list their
// partners through visit(row, emit)); lower parts are filled through atomic cursors and sorted.
// With dedupe, repeated partners are merged and the rows compacted.
template <class V>
int build_symmetric(sa_graph* g, int64_t n, int threads, int32_t kind, uint64_t seed, bool dedupe, V visit) {
  std::vector<std::atomic<int64_t>> deg(n);
  parallel_chunks(n, 1 << 14, threads, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) deg[i].store(0, std::memory_order_relaxed);
  });
  parallel_chunks(n, 4096, threads, [&](int64_t lo, int64_t hi) {
    for (int64_t r = lo; r < hi; ++r)
      visit(r, [&](int64_t j) {
        deg[r].fetch_add(1, std::memory_order_relaxed);
        deg[j].fetch_add(1, std::memory_order_relaxed);
      });
  });
  std::vector<int64_t> ptr(n + 1);
  ptr[0] = 0;
  for (int64_t i = 0; i < n; ++i) ptr[i + 1] = ptr[i] + deg[i].load(std::memory_order_relaxed);
  const int64_t total = ptr[n];
  std::vector<int32_t> idx((size_t)total);
  std::vector<double> val((size_t)total);
  parallel_chunks(n, 1 << 14, threads, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) deg[i].store(ptr[i], std::memory_order_relaxed);  // cursors
  });
  parallel_chunks(n, 4096, threads, [&](int64_t lo, int64_t hi) {
    for (int64_t r = lo; r < hi; ++r)
      visit(r, [&](int64_t j) {
        const double v = coupling(kind, seed, r, j);
        const int64_t a = deg[r].fetch_add(1, std::memory_order_relaxed);
        idx[a] = (int32_t)j;
        val[a] = v;
        const int64_t b = deg[j].fetch_add(1, std::memory_order_relaxed);
        idx[b] = (int32_t)r;
        val[b] = v;
      });
  });
  // sort each row by column (the values of an edge are a function of the pair, so ties are equal)
  std::vector<int64_t> keep(n, 0);
  parallel_chunks(n, 4096, threads, [&](int64_t lo, int64_t hi) {
    std::vector<std::pair<int32_t, double>> row;
    for (int64_t i = lo; i < hi; ++i) {
      const int64_t a = ptr[i], b = ptr[i + 1];
      row.clear();
      for (int64_t e = a; e < b; ++e) row.emplace_back(idx[e], val[e]);
      std::sort(row.begin(), row.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
      int64_t w = a;
      for (size_t q = 0; q < row.size(); ++q) {
        if (dedupe && q > 0 && row[q].first == row[q - 1].first) continue;
        idx[w] = row[q].first;
        val[w] = row[q].second;
        ++w;
      }
      keep[i] = w - a;
    }
  });
  g->n = n;
  g->indptr.assign(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) g->indptr[i + 1] = g->indptr[i] + keep[i];
  if (g->indptr[n] == total) {
    g->indices.swap(idx);
    g->vals.swap(val);
  } else {
    g->indices.resize((size_t)g->indptr[n]);
    g->vals.resize((size_t)g->indptr[n]);
    parallel_chunks(n, 4096, threads, [&](int64_t lo, int64_t hi) {
      for (int64_t i = lo; i < hi; ++i) {
        std::copy(idx.begin() + ptr[i], idx.begin() + ptr[i] + keep[i], g->indices.begin() + g->indptr[i]);
        std::copy(val.begin() + ptr[i], val.begin() + ptr[i] + keep[i], g->vals.begin() + g->indptr[i]);
      }
    });
  }
  // drop isolated nodes (no couplings), renumbering the others in order, as the instance loader does
  int64_t isolated = 0;
  for (int64_t i = 0; i < n; ++i) isolated += keep[i] == 0;
  if (isolated) {
    std::vector<int32_t> rename(n, -1);
    int64_t next = 0;
    for (int64_t i = 0; i < n; ++i)
      if (keep[i]) rename[i] = (int32_t)next++;
    std::vector<int64_t> ip(next + 1, 0);
    for (int64_t i = 0, q = 0; i < n; ++i)
      if (keep[i]) ip[q + 1] = ip[q] + keep[i], ++q;
    parallel_chunks(g->indptr[n], 1 << 16, threads, [&](int64_t lo, int64_t hi) {
      for (int64_t e = lo; e < hi; ++e) g->indices[e] = rename[g->indices[e]];
    });
    g->indptr.swap(ip);
    g->n = next;
  }
  if (g->indptr[g->n] >= (int64_t)1 << 31) {
    set_error("graph has 2^31 or more stored couplings: int32 CSR indices cannot address them");
    return SA_E_ARG;
  }
  return SA_OK;
}

int resolve_threads(int32_t threads) {
  if (threads > 0) return threads;
  const unsigned hw = std::thread::hardware_concurrency();
  return hw ? (int)hw : 1;
}

}  // namespace
}  // namespace sa

using namespace sa;

extern "C" {

int sa_graph_erdos_renyi(int64_t n, double p, uint64_t seed, int32_t coupling_kind, int32_t threads,
                         sa_graph** out) {
  SA_REQUIRE(out && n >= 2 && n < ((int64_t)1 << 31), "need 2 <= n < 2^31 and an output handle");
  SA_REQUIRE(p >= 0.0 && p <= 1.0, "p must be in [0, 1]");
  SA_REQUIRE(coupling_kind == 0 || coupling_kind == 1, "coupling_kind: 0 = all -1, 1 = +-1");
  *out = nullptr;
  sa_graph* g = new (std::nothrow) sa_graph;
  if (!g) return SA_E_OOM;
  const double lq = std::log1p(-std::min(p, 1.0 - 1e-16));
  try {
    const int rc = build_symmetric(g, n, resolve_threads(threads), coupling_kind, seed, false,
                                   [&](int64_t r, auto emit) { er_row(n, p, lq, seed, r, emit); });
    if (rc) {
      delete g;
      return rc;
    }
  } catch (const std::bad_alloc&) {
    delete g;
    set_error("sa_graph_erdos_renyi: out of host memory");
    return SA_E_OOM;
  }
  *out = g;
  return SA_OK;
}

int sa_graph_barabasi_albert(int64_t n, int32_t m, uint64_t seed, int32_t coupling_kind, int32_t threads,
                             sa_graph** out) {
  SA_REQUIRE(out && n >= 2 && n < ((int64_t)1 << 31), "need 2 <= n < 2^31 and an output handle");
  SA_REQUIRE(m >= 1 && (int64_t)m * n < ((int64_t)1 << 40), "m must be >= 1");
  SA_REQUIRE(coupling_kind == 0 || coupling_kind == 1, "coupling_kind: 0 = all -1, 1 = +-1");
  *out = nullptr;
  sa_graph* g = new (std::nothrow) sa_graph;
  if (!g) return SA_E_OOM;
  try {
    // edge k joins its source v = k / m and an older node (its target), listed by the source's row
    const int rc = build_symmetric(g, n, resolve_threads(threads), coupling_kind, seed, true,
                                   [&](int64_t v, auto emit) {
                                     for (int32_t i = 0; i < m; ++i) {
                                       const int64_t t = ba_target(v * m + i, m, seed);
                                       if (t != v) emit(t);
                                     }
                                   });
    if (rc) {
      delete g;
      return rc;
    }
  } catch (const std::bad_alloc&) {
    delete g;
    set_error("sa_graph_barabasi_albert: out of host memory");
    return SA_E_OOM;
  }
  *out = g;
  return SA_OK;
}

int64_t sa_graph_nnz(const sa_graph* g) { return g ? g->indptr.back() : -1; }

int64_t sa_graph_n(const sa_graph* g) { return g ? g->n : -1; }

int sa_graph_export(const sa_graph* g, int64_t* indptr, int32_t* indices, double* vals) {
  SA_REQUIRE(g && indptr && indices && vals, "bad arguments");
  std::copy(g->indptr.begin(), g->indptr.end(), indptr);
  std::copy(g->indices.begin(), g->indices.end(), indices);
  std::copy(g->vals.begin(), g->vals.end(), vals);
  return SA_OK;
}

void sa_graph_free(sa_graph* g) { delete g; }

}  // extern "C"
