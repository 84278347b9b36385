// Seeded synthetic RC forests and timing DAGs, plus the host-side fanin CSR
// build.  Every random value is drawn from a counter-based stream keyed by
// (seed, object index, purpose), so output is independent of thread count.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/otg_gen.h"
#include "../../include/otg_prop.h"

namespace {

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// splitmix64 stream seeded from (seed, index, purpose)
struct Rng {
  uint64_t s;
  Rng(uint64_t seed, uint64_t idx, uint64_t purpose)
      : s(mix64(seed ^ mix64(idx * 0x9e3779b97f4a7c15ULL + purpose))) {}
  uint64_t next() {
    s += 0x9e3779b97f4a7c15ULL;
    return mix64(s);
  }
  double u01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  // uniform integer in [lo, hi]
  int64_t range(int64_t lo, int64_t hi) {
    uint64_t span = (uint64_t)(hi - lo) + 1;
    return lo + (int64_t)(next() % span);
  }
  double loguniform(double a, double b) {
    if (a <= 0 || b <= 0) return a;
    return std::exp(std::log(a) + u01() * (std::log(b) - std::log(a)));
  }
};

int resolve_threads(int n) {
  if (n > 0) return n;
  unsigned h = std::thread::hardware_concurrency();
  return h ? (int)std::min(h, 64u) : 1;
}

template <class F>
void parallel_ranges(int64_t total, int nthreads, F f) {
  nthreads = resolve_threads(nthreads);
  if (nthreads <= 1 || total < 1024) {
    f((int64_t)0, total);
    return;
  }
  std::atomic<int64_t> next{0};
  const int64_t chunk = std::max<int64_t>(256, total / (nthreads * 16));
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t)
    th.emplace_back([&] {
      for (;;) {
        int64_t b = next.fetch_add(chunk);
        if (b >= total) break;
        f(b, std::min(total, b + chunk));
      }
    });
  for (auto& t : th) t.join();
}

int pick_class(const otg_gen_rc_params* p, Rng& r) {
  double tot = 0;
  for (int i = 0; i < p->num_classes; ++i) tot += p->classes[i].weight;
  double x = r.u01() * tot;
  for (int i = 0; i < p->num_classes; ++i) {
    if (x < p->classes[i].weight) return i;
    x -= p->classes[i].weight;
  }
  return p->num_classes - 1;
}

bool valid_params(const otg_gen_rc_params* p) {
  if (!p || p->num_nets < 0 || p->num_classes < 1 || p->num_classes > 4 || p->first_net < 0)
    return false;
  for (int i = 0; i < p->num_classes; ++i) {
    const auto& c = p->classes[i];
    if (c.min_nodes < 1 || c.max_nodes < c.min_nodes || c.weight < 0) return false;
    if (c.shape < 0 || c.shape > 2) return false;
  }
  return true;
}

}  // namespace

extern "C" otg_status otg_gen_rc_sizes(const otg_gen_rc_params* p, int32_t* nodes_start,
                                       int64_t* total_nodes) {
  if (!valid_params(p) || !nodes_start) return OTG_ERR_INVALID_ARGUMENT;
  int64_t tot = 0;
  nodes_start[0] = 0;
  for (int32_t k = 0; k < p->num_nets; ++k) {
    Rng r(p->seed, (uint64_t)(p->first_net + k), 1);
    const auto& c = p->classes[pick_class(p, r)];
    int64_t n = r.range(c.min_nodes, c.max_nodes);
    tot += n;
    if (tot > INT32_MAX) return OTG_ERR_INVALID_ARGUMENT;
    nodes_start[k + 1] = (int32_t)tot;
  }
  if (total_nodes) *total_nodes = tot;
  return OTG_OK;
}

extern "C" otg_status otg_gen_rc_fill(const otg_gen_rc_params* p, const int32_t* nodes_start,
                                      int32_t* edges, int32_t* roots, float* edge_res,
                                      float* node_cap, float* slew_in, int32_t* parent_gen,
                                      int nthreads) {
  if (!valid_params(p) || !nodes_start || !edges || !roots || !edge_res || !node_cap)
    return OTG_ERR_INVALID_ARGUMENT;
  parallel_ranges(p->num_nets, nthreads, [&](int64_t b, int64_t e) {
    std::vector<int32_t> par, perm;
    std::vector<uint8_t> leaf;
    for (int64_t k = b; k < e; ++k) {
      Rng rc(p->seed, (uint64_t)(p->first_net + k), 1);
      const auto& cls = p->classes[pick_class(p, rc)];
      const int64_t off = nodes_start[k];
      const int32_t n = nodes_start[k + 1] - nodes_start[k];
      if (n <= 0) { roots[k] = 0; continue; }
      Rng r(p->seed, (uint64_t)(p->first_net + k), 2);
      par.assign(n, -1);
      leaf.assign(n, 1);
      for (int32_t i = 1; i < n; ++i) {
        int32_t q;
        if (cls.shape == OTG_SHAPE_BALANCED) q = (i - 1) / 2;
        else if (cls.shape == OTG_SHAPE_CHAIN && r.u01() >= cls.branch_prob) q = i - 1;
        else q = (int32_t)r.range(0, i - 1);
        par[i] = q;
        leaf[q] = 0;
      }
      // node numbering: identity or a random permutation (driver moves too)
      perm.resize(n);
      for (int32_t i = 0; i < n; ++i) perm[i] = i;
      if (p->relabel)
        for (int32_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[r.range(0, i)]);
      roots[k] = perm[0];
      if (parent_gen)
        for (int32_t i = 0; i < n; ++i) parent_gen[off + i] = par[i];
      const int64_t eoff = off - k;
      for (int32_t i = 1; i < n; ++i) {
        int32_t s = perm[par[i]], t = perm[i];
        if (p->random_orient && (r.next() & 1)) std::swap(s, t);
        edges[2 * (eoff + i - 1)] = s;
        edges[2 * (eoff + i - 1) + 1] = t;
        edge_res[eoff + i - 1] = (float)(r.loguniform(p->r_min, p->r_max) * p->r_unit);
      }
      for (int32_t i = 0; i < n; ++i) {
        double g = r.loguniform(p->c_min, p->c_max);
        double pin = (i > 0 && leaf[i] && p->pin_max > 0) ? r.loguniform(p->pin_min, p->pin_max) : 0.0;
        for (int c = 0; c < 4; ++c)
          node_cap[(off + perm[i]) * 4 + c] = (float)(g + pin * p->corner_scale[c]);
      }
      if (slew_in) {
        double s = p->slew_min + r.u01() * (p->slew_max - p->slew_min);
        for (int c = 0; c < 4; ++c) slew_in[k * 4 + c] = (float)s;
      }
    }
  });
  return OTG_OK;
}

extern "C" otg_status otg_gen_dag(const otg_gen_dag_params* p, int32_t* arc_from,
                                  int32_t* arc_to, int nthreads) {
  if (!p || p->num_pins < 2 || p->num_layers < 2 || p->num_layers > p->num_pins ||
      p->fanout_cap < 1 || !arc_from || !arc_to || p->num_arcs < 0)
    return OTG_ERR_INVALID_ARGUMENT;
  const int32_t n = p->num_pins, L = p->num_layers;
  auto layer_begin = [&](int32_t l) { return (int32_t)((int64_t)n * l / L); };
  const int32_t src_pins = layer_begin(L - 1);  // pins of layers 0..L-2 have fanout
  if (p->num_arcs < src_pins || p->num_arcs > (int64_t)src_pins * p->fanout_cap)
    return OTG_ERR_INVALID_ARGUMENT;
  // discrete power law CDF over 1..cap
  std::vector<double> cdf(p->fanout_cap);
  double acc = 0;
  for (int k = 1; k <= p->fanout_cap; ++k) cdf[k - 1] = (acc += std::pow((double)k, -p->fanout_alpha));
  for (auto& x : cdf) x /= acc;
  std::vector<int32_t> deg(src_pins);
  parallel_ranges(src_pins, nthreads, [&](int64_t b, int64_t e) {
    for (int64_t u = b; u < e; ++u) {
      Rng r(p->seed, (uint64_t)u, 11);
      double x = r.u01();
      deg[u] = (int32_t)(std::lower_bound(cdf.begin(), cdf.end(), x) - cdf.begin()) + 1;
    }
  });
  // adjust to the exact arc count with a single seeded stream
  int64_t tot = 0;
  for (auto d : deg) tot += d;
  Rng adj(p->seed, 0, 12);
  while (tot != p->num_arcs) {
    int32_t u = (int32_t)adj.range(0, src_pins - 1);
    if (tot > p->num_arcs && deg[u] > 1) { --deg[u]; --tot; }
    else if (tot < p->num_arcs && deg[u] < p->fanout_cap) { ++deg[u]; ++tot; }
  }
  std::vector<int64_t> off(src_pins + 1, 0);
  for (int32_t u = 0; u < src_pins; ++u) off[u + 1] = off[u] + deg[u];
  // pin numbering
  std::vector<int32_t> id(n);
  for (int32_t i = 0; i < n; ++i) id[i] = i;
  if (p->shuffle_ids) {
    Rng r(p->seed, 0, 13);
    for (int32_t i = n - 1; i > 0; --i) std::swap(id[i], id[r.range(0, i)]);
  }
  parallel_ranges(src_pins, nthreads, [&](int64_t b, int64_t e) {
    int32_t l = 0;
    for (int64_t u = b; u < e; ++u) {
      while (layer_begin(l + 1) <= u) ++l;
      Rng r(p->seed, (uint64_t)u, 14);
      for (int64_t j = off[u]; j < off[u + 1]; ++j) {
        int32_t tl = l + 1;
        if (j > off[u] && p->skip_prob > 0 && r.u01() < p->skip_prob)
          tl = std::min<int32_t>(L - 1, l + (int32_t)r.range(p->skip_min, p->skip_max));
        int32_t lb = layer_begin(tl), le = layer_begin(tl + 1);
        int32_t v = (int32_t)r.range(lb, le - 1);
        arc_from[j] = id[u];
        arc_to[j] = id[v];
      }
    }
  });
  return OTG_OK;
}

extern "C" otg_status otg_build_fanin_csr(int32_t n, int64_t m, const int32_t* arc_from,
                                          const int32_t* arc_to, const uint8_t* cand,
                                          const uint8_t* skip_arc, int32_t* fanin_start,
                                          int32_t* fanin, int32_t* out_degree,
                                          int32_t* frontier, int32_t* first_size,
                                          int64_t* num_kept) {
  if (n < 0 || m < 0 || !fanin_start || !fanin || !out_degree || !frontier || !first_size ||
      (m > 0 && (!arc_from || !arc_to)))
    return OTG_ERR_INVALID_ARGUMENT;
  for (int64_t j = 0; j < m; ++j)
    if (arc_from[j] < 0 || arc_from[j] >= n || arc_to[j] < 0 || arc_to[j] >= n)
      return OTG_ERR_INVALID_ARGUMENT;
  auto keep = [&](int64_t j) {
    if (skip_arc && skip_arc[j]) return false;
    if (cand && (!cand[arc_from[j]] || !cand[arc_to[j]])) return false;
    return true;
  };
  std::memset(fanin_start, 0, sizeof(int32_t) * ((size_t)n + 1));
  std::memset(out_degree, 0, sizeof(int32_t) * (size_t)n);
  int64_t kept = 0;
  for (int64_t j = 0; j < m; ++j)
    if (keep(j)) {
      fanin_start[arc_to[j] + 1]++;
      out_degree[arc_from[j]]++;
      ++kept;
    }
  if (kept > INT32_MAX) return OTG_ERR_INVALID_ARGUMENT;
  for (int32_t i = 0; i < n; ++i) fanin_start[i + 1] += fanin_start[i];
  std::vector<int32_t> fill(fanin_start, fanin_start + n);
  for (int64_t j = 0; j < m; ++j)
    if (keep(j)) fanin[fill[arc_to[j]]++] = arc_from[j];
  int32_t f = 0;
  for (int32_t i = 0; i < n; ++i)
    if ((!cand || cand[i]) && out_degree[i] == 0) frontier[f++] = i;
  *first_size = f;
  if (num_kept) *num_kept = kept;
  return OTG_OK;
}

extern "C" otg_status otg_build_fanout_csr(int32_t num_pins, const int32_t* fanin_start,
                                           const int32_t* fanin, int32_t* fanout_start,
                                           int32_t* fanout_arc, int32_t* arc_to) {
  if (num_pins < 0 || !fanin_start || !fanout_start) return OTG_ERR_INVALID_ARGUMENT;
  const int64_t m = fanin_start[num_pins];
  if (m > 0 && (!fanin || !fanout_arc || !arc_to)) return OTG_ERR_INVALID_ARGUMENT;
  for (int32_t v = 0; v <= num_pins; ++v) fanout_start[v] = 0;
  for (int32_t v = 0; v < num_pins; ++v) {
    if (fanin_start[v + 1] < fanin_start[v]) return OTG_ERR_INVALID_ARGUMENT;
    for (int32_t j = fanin_start[v]; j < fanin_start[v + 1]; ++j) {
      const int32_t u = fanin[j];
      if (u < 0 || u >= num_pins) return OTG_ERR_INVALID_ARGUMENT;
      arc_to[j] = v;
      ++fanout_start[u + 1];
    }
  }
  for (int32_t v = 0; v < num_pins; ++v) fanout_start[v + 1] += fanout_start[v];
  std::vector<int32_t> next(fanout_start, fanout_start + num_pins);
  for (int64_t j = 0; j < m; ++j) fanout_arc[next[fanin[j]]++] = (int32_t)j;
  return OTG_OK;
}
