// Host lattice octree, U/V interaction lists and Morton keys.
//
// The tree is a cubic lattice octree over the panel centroids, ordered by the
// reference's Morton key (fmm3d/geometry.py:736-799) evaluated in FP64 at level
// 21 inside the root cube.  Because boxes sit on a lattice, every M2L
// translation depends only on an integer offset between same-level boxes, so
// translation operators are shared tables and M2L becomes a batched GEMM.
//
// Lists (ws = well-separation radius in boxes):
//   adjacent(A,B): same level and max|coord(A)-coord(B)| <= ws
//   V(T): children of non-leaf nodes adjacent to parent(T) that are not
//         adjacent to T (M2L sources of T)
//   U(T): for a leaf T, the leaves S with adjacent(anc_m(T), anc_m(S)) at
//         m = min(level T, level S) (P2P sources of T)
// These cover every ordered pair of points exactly once.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>

#include "ed_internal.h"

namespace ed {

// Spreads the low 21 bits of v to every third bit (bit b -> bit 3b) with the
// shift-and-mask ("magic numbers") method: five steps of doubling shifts.
static inline uint64_t spread3(uint64_t v) {
  v &= 0x1fffffULL;
  v = (v | (v << 32)) & 0x1f00000000ffffULL;
  v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
  v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
  v = (v | (v << 2)) & 0x1249249249249249ULL;
  return v;
}

// Bit b of x, y, z goes to bit 3b, 3b+1, 3b+2 of the key, b < bits (<= 21).
static inline uint64_t interleave3(uint64_t x, uint64_t y, uint64_t z, int bits) {
  const uint64_t mask = bits >= 21 ? 0x1fffffULL : ((1ULL << bits) - 1);
  return spread3(x & mask) | (spread3(y & mask) << 1) | (spread3(z & mask) << 2);
}

// fmm3d/geometry.py:696-703 and :785-799 in FP64: normalised = (p - min)/extent
// clamped to [0, 1-1e-9], extent = max(max-min, eps); floor(x 2^level) clamped.
static inline void quantize(const double* p, const double* mins, const double* ext,
                            int level, int64_t* ijk) {
  const double hi = 1.0 - 1e-9;
  const double scale = (double)(1LL << level);
  const int64_t maxi = (1LL << level) - 1;
  for (int a = 0; a < 3; ++a) {
    double v = (p[a] - mins[a]) / ext[a];
    v = v < 0.0 ? 0.0 : v;
    v = v > hi ? hi : v;
    int64_t q = (int64_t)std::floor(v * scale);
    ijk[a] = q < 0 ? 0 : (q > maxi ? maxi : q);
  }
}

void morton_keys_host(const double* pts, int64_t n, int level,
                      const double* bbox_center, const double* bbox_half,
                      int64_t* keys) {
  if (n <= 0) return;
  double mins[3], maxs[3], ext[3];
  if (bbox_center) {
    for (int a = 0; a < 3; ++a) {
      mins[a] = bbox_center[a] - bbox_half[a];
      maxs[a] = bbox_center[a] + bbox_half[a];
    }
  } else {
    // geometry.py:678-693: tight bbox, degenerate axes padded by
    // 1e-12 * (1 + max_extent).
    for (int a = 0; a < 3; ++a) { mins[a] = pts[a]; maxs[a] = pts[a]; }
    for (int64_t i = 1; i < n; ++i)
      for (int a = 0; a < 3; ++a) {
        mins[a] = std::min(mins[a], pts[3 * i + a]);
        maxs[a] = std::max(maxs[a], pts[3 * i + a]);
      }
    double e[3], me = 0.0;
    for (int a = 0; a < 3; ++a) { e[a] = maxs[a] - mins[a]; me = std::max(me, std::fabs(e[a])); }
    const double padded = 1e-12 * (1.0 + me);
    for (int a = 0; a < 3; ++a) maxs[a] = mins[a] + (e[a] > 0 ? e[a] : padded);
  }
  const double eps = 2.220446049250313e-16;
  for (int a = 0; a < 3; ++a) ext[a] = std::max(maxs[a] - mins[a], eps);
  // each key depends on its point alone: contiguous ranges on worker threads
  auto run = [&](int64_t i0, int64_t i1) {
    for (int64_t i = i0; i < i1; ++i) {
      int64_t ijk[3];
      quantize(pts + 3 * i, mins, ext, level, ijk);
      keys[i] = (int64_t)interleave3(ijk[0], ijk[1], ijk[2], level);
    }
  };
  const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
  const int nthr = (int)std::max<int64_t>(1, std::min<int64_t>(std::min(hw, 8), n / 8192));
  std::vector<std::thread> pool;
  for (int k = 1; k < nthr; ++k) pool.emplace_back(run, n * k / nthr, n * (k + 1) / nthr);
  run(0, n / nthr);
  for (auto& th : pool) th.join();
}

int64_t HostTree::n_leaves() const {
  int64_t c = 0;
  for (auto v : is_leaf) c += v;
  return c;
}

static constexpr int kKeyLevel = 21;

int build_host_tree(const double* C, int64_t n, int leaf_size, int max_level,
                    int ws, HostTree& t) {
  if (n <= 0) return fail(ED_ERR_ARG, "tree build: n must be positive");
  if (leaf_size < 1) return fail(ED_ERR_ARG, "tree build: leaf_size must be >= 1");
  if (max_level < 1 || max_level > 20)
    return fail(ED_ERR_ARG, "tree build: max_level must be in [1, 20]");
  if (ws < 1 || ws > 2) return fail(ED_ERR_ARG, "tree build: ws must be 1 or 2");
  t = HostTree();
  t.n = n;
  t.leaf_size = leaf_size;
  t.max_level = max_level;
  t.ws = ws;
  double mins[3], maxs[3];
  for (int a = 0; a < 3; ++a) { mins[a] = C[a]; maxs[a] = C[a]; }
  for (int64_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      const double v = C[3 * i + a];
      if (!std::isfinite(v)) return fail(ED_ERR_ARG, "tree build: non-finite centroid");
      mins[a] = std::min(mins[a], v);
      maxs[a] = std::max(maxs[a], v);
    }
  double half = 0.0;
  for (int a = 0; a < 3; ++a) {
    t.center[a] = 0.5 * (mins[a] + maxs[a]);
    half = std::max(half, 0.5 * (maxs[a] - mins[a]));
  }
  half = half * (1.0 + std::ldexp(1.0, -20)) + 1e-300;
  t.half = half;
  const double bh[3] = {half, half, half};
  std::vector<int64_t> keys(n);
  morton_keys_host(C, n, kKeyLevel, t.center, bh, keys.data());
  // stable LSD radix sort of (key, index) on 11-bit digits (the 2,048
  // counters stay in L1), skipping a digit every key shares: key order, ties
  // in index order
  {
    std::vector<uint64_t> ka(n), kb(n);
    std::vector<int64_t> ia(n), ib(n);
    for (int64_t i = 0; i < n; ++i) { ka[i] = (uint64_t)keys[i]; ia[i] = i; }
    constexpr int kBits = 11, kBuckets = 1 << kBits;
    std::vector<int64_t> cnt(kBuckets + 1);
    for (int sh = 0; sh < 3 * kKeyLevel; sh += kBits) {
      std::fill(cnt.begin(), cnt.end(), 0);
      for (int64_t i = 0; i < n; ++i) ++cnt[((ka[i] >> sh) & (kBuckets - 1)) + 1];
      if (cnt[((ka[0] >> sh) & (kBuckets - 1)) + 1] == n) continue;
      for (int d = 0; d < kBuckets; ++d) cnt[d + 1] += cnt[d];
      for (int64_t i = 0; i < n; ++i) {
        const int64_t dst = cnt[(ka[i] >> sh) & (kBuckets - 1)]++;
        kb[dst] = ka[i];
        ib[dst] = ia[i];
      }
      ka.swap(kb);
      ia.swap(ib);
    }
    t.perm = std::move(ia);
    t.keys = std::move(ka);
  }

  // Level-by-level node construction.
  struct Pending { int64_t b, e, parent; };
  std::vector<Pending> frontier{{0, n, -1}};
  int lvl = 0;
  while (!frontier.empty()) {
    const int64_t base = (int64_t)t.level.size();
    t.level_begin.push_back(base);
    for (auto& f : frontier) {
      t.level.push_back(lvl);
      t.begin.push_back(f.b);
      t.end.push_back(f.e);
      t.parent.push_back(f.parent);
      const uint64_t pref = t.keys[f.b] >> (3 * (kKeyLevel - lvl));
      // de-interleave the level-lvl prefix into integer coords
      int32_t xyz[3] = {0, 0, 0};
      for (int b = 0; b < lvl; ++b)
        for (int a = 0; a < 3; ++a) xyz[a] |= (int32_t)((pref >> (3 * b + a)) & 1ULL) << b;
      t.coord.insert(t.coord.end(), xyz, xyz + 3);
      const bool leaf = (f.e - f.b) <= leaf_size || lvl >= max_level;
      t.is_leaf.push_back(leaf ? 1 : 0);
      t.first_child.push_back(-1);
      t.n_child.push_back(0);
    }
    std::vector<Pending> next;
    for (size_t k = 0; k < frontier.size(); ++k) {
      const int64_t idx = base + (int64_t)k;
      if (t.is_leaf[idx]) continue;
      const int shift = 3 * (kKeyLevel - lvl - 1);
      t.first_child[idx] = (int32_t)(base + (int64_t)frontier.size() + (int64_t)next.size());
      int64_t s0 = frontier[k].b;
      int nc = 0;
      for (int64_t i = frontier[k].b + 1; i <= frontier[k].e; ++i) {
        if (i == frontier[k].e || (t.keys[i] >> shift) != (t.keys[i - 1] >> shift)) {
          next.push_back({s0, i, idx});
          s0 = i;
          ++nc;
        }
      }
      t.n_child[idx] = nc;
    }
    frontier.swap(next);
    ++lvl;
  }
  t.depth = lvl - 1;
  t.level_begin.push_back((int64_t)t.level.size());
  const int64_t M = (int64_t)t.level.size();
  t.node_c.resize(3 * M);
  t.node_s.resize(M);
  for (int64_t i = 0; i < M; ++i) {
    const double side = 2.0 * half / std::ldexp(1.0, t.level[i]);
    t.node_s[i] = side;
    for (int a = 0; a < 3; ++a) t.node_c[3 * i + a] = -half + (t.coord[3 * i + a] + 0.5) * side;
  }
  t.geom_center.assign(t.center, t.center + 3);
  t.geom_pt_off = {0, n};

  // Locational-code hash: (1 << 3l) | prefix -> node, open addressing with
  // linear probing (codes are never 0, so 0 marks an empty slot).
  auto code_of = [&](int l, int64_t x, int64_t y, int64_t z) -> uint64_t {
    return (1ULL << (3 * l)) | interleave3((uint64_t)x, (uint64_t)y, (uint64_t)z, l);
  };
  int hbits = 4;
  while ((1LL << hbits) < 4 * M) ++hbits;
  const uint64_t hmask = (1ULL << hbits) - 1;
  std::vector<uint64_t> hkey((size_t)1 << hbits, 0);
  std::vector<int64_t> hval((size_t)1 << hbits, -1);
  auto slot_of = [&](uint64_t code) {
    return (size_t)((code * 0x9E3779B97F4A7C15ULL) >> (64 - hbits));
  };
  for (int64_t i = 0; i < M; ++i) {
    const uint64_t c = code_of(t.level[i], t.coord[3 * i], t.coord[3 * i + 1], t.coord[3 * i + 2]);
    size_t h = slot_of(c);
    while (hkey[h] != 0 && hkey[h] != c) h = (h + 1) & hmask;
    hkey[h] = c;
    hval[h] = i;
  }
  auto find_node = [&](uint64_t c) -> int64_t {
    size_t h = slot_of(c);
    while (hkey[h] != 0) {
      if (hkey[h] == c) return hval[h];
      h = (h + 1) & hmask;
    }
    return -1;
  };

  auto colleagues = [&](int64_t a, std::vector<int64_t>& out) {
    out.clear();
    const int l = t.level[a];
    const int64_t lim = 1LL << l;
    for (int dz = -ws; dz <= ws; ++dz)
      for (int dy = -ws; dy <= ws; ++dy)
        for (int dx = -ws; dx <= ws; ++dx) {
          const int64_t x = t.coord[3 * a] + dx, y = t.coord[3 * a + 1] + dy,
                        z = t.coord[3 * a + 2] + dz;
          if (x < 0 || y < 0 || z < 0 || x >= lim || y >= lim || z >= lim) continue;
          const int64_t f = find_node(code_of(l, x, y, z));
          if (f >= 0) out.push_back(f);
        }
  };
  auto adjacent = [&](int64_t a, int64_t b) {
    for (int k = 0; k < 3; ++k)
      if (std::abs(t.coord[3 * a + k] - t.coord[3 * b + k]) > ws) return false;
    return true;
  };

  // V and U lists: nodes are independent, so contiguous node ranges go to
  // worker threads (read-only access to the tree and the hash); each range
  // fills its own lists and the ranges are joined in node order, so the
  // result does not depend on the thread count.
  struct Lists {
    std::vector<int64_t> v_src, u_src, v_cnt, u_cnt;
    std::vector<int32_t> v_off;
  };
  auto build_lists = [&](int64_t n0, int64_t n1, Lists& out) {
    std::vector<int64_t> col, col2, stack, us;
    std::vector<std::pair<int64_t, int32_t>> vs;
    out.v_cnt.assign((size_t)(n1 - n0), 0);
    out.u_cnt.assign((size_t)(n1 - n0), 0);
    for (int64_t nidx = n0; nidx < n1; ++nidx) {
      vs.clear();
      const int64_t P = t.parent[nidx];
      if (P >= 0) {
        colleagues(P, col);
        for (int64_t c : col) {
          if (t.is_leaf[c]) continue;
          for (int64_t s = t.first_child[c]; s < t.first_child[c] + t.n_child[c]; ++s) {
            if (adjacent(s, nidx)) continue;
            const int ox = t.coord[3 * nidx] - t.coord[3 * s];
            const int oy = t.coord[3 * nidx + 1] - t.coord[3 * s + 1];
            const int oz = t.coord[3 * nidx + 2] - t.coord[3 * s + 2];
            vs.emplace_back(s, offset_slot(ws, ox, oy, oz));
          }
        }
        std::sort(vs.begin(), vs.end());
      }
      for (auto& e : vs) { out.v_src.push_back(e.first); out.v_off.push_back(e.second); }
      out.v_cnt[(size_t)(nidx - n0)] = (int64_t)vs.size();

      us.clear();
      if (t.is_leaf[nidx]) {
        for (int64_t a = nidx; a >= 0; a = t.parent[a]) {
          colleagues(a, col2);
          for (int64_t c : col2) {
            if (t.is_leaf[c]) {
              us.push_back(c);
            } else if (a == nidx) {
              stack.assign(1, c);
              while (!stack.empty()) {
                const int64_t d = stack.back();
                stack.pop_back();
                if (t.is_leaf[d]) { us.push_back(d); continue; }
                for (int64_t s = t.first_child[d]; s < t.first_child[d] + t.n_child[d]; ++s)
                  stack.push_back(s);
              }
            }
          }
        }
        std::sort(us.begin(), us.end());
        us.erase(std::unique(us.begin(), us.end()), us.end());
      }
      out.u_src.insert(out.u_src.end(), us.begin(), us.end());
      out.u_cnt[(size_t)(nidx - n0)] = (int64_t)us.size();
    }
  };
  const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
  const int nthr = (int)std::max<int64_t>(1, std::min<int64_t>(std::min(hw, 8), M / 128));
  std::vector<Lists> parts((size_t)nthr);
  {
    std::vector<std::thread> pool;
    for (int k = 1; k < nthr; ++k)
      pool.emplace_back(build_lists, M * k / nthr, M * (k + 1) / nthr, std::ref(parts[(size_t)k]));
    build_lists(0, M / nthr, parts[0]);
    for (auto& th : pool) th.join();
  }
  t.v_ptr.assign(M + 1, 0);
  t.u_ptr.assign(M + 1, 0);
  int64_t nidx = 0;
  for (auto& L : parts) {
    t.v_src.insert(t.v_src.end(), L.v_src.begin(), L.v_src.end());
    t.v_off.insert(t.v_off.end(), L.v_off.begin(), L.v_off.end());
    t.u_src.insert(t.u_src.end(), L.u_src.begin(), L.u_src.end());
    for (size_t q = 0; q < L.v_cnt.size(); ++q, ++nidx) {
      t.v_ptr[nidx + 1] = t.v_ptr[nidx] + L.v_cnt[q];
      t.u_ptr[nidx + 1] = t.u_ptr[nidx] + L.u_cnt[q];
    }
  }
  return ED_OK;
}

int merge_host_trees(std::vector<HostTree>& parts, HostTree& out) {
  if (parts.empty()) return fail(ED_ERR_ARG, "merge: no geometries");
  if (parts.size() == 1) { out = std::move(parts[0]); return ED_OK; }
  out = HostTree();
  const size_t G = parts.size();
  out.leaf_size = parts[0].leaf_size;
  out.max_level = parts[0].max_level;
  out.ws = parts[0].ws;
  std::vector<int64_t> pt_off(G + 1, 0);
  int depth = 0;
  for (size_t g = 0; g < G; ++g) {
    pt_off[g + 1] = pt_off[g] + parts[g].n;
    depth = std::max(depth, parts[g].depth);
  }
  out.n = pt_off[G];
  out.depth = depth;
  out.geom_pt_off = pt_off;
  // new node index: level-major, then geometry, then the part's own order
  std::vector<std::vector<int64_t>> nid(G);
  int64_t next = 0;
  for (size_t g = 0; g < G; ++g) nid[g].assign(parts[g].level.size(), -1);
  for (int l = 0; l <= depth; ++l) {
    out.level_begin.push_back(next);
    for (size_t g = 0; g < G; ++g) {
      const HostTree& t = parts[g];
      if (l > t.depth) continue;
      for (int64_t i = t.level_begin[l]; i < t.level_begin[l + 1]; ++i) nid[g][i] = next++;
    }
  }
  out.level_begin.push_back(next);
  const int64_t M = next;
  out.level.assign(M, 0);
  out.begin.assign(M, 0);
  out.end.assign(M, 0);
  out.parent.assign(M, -1);
  out.first_child.assign(M, -1);
  out.n_child.assign(M, 0);
  out.coord.assign(3 * M, 0);
  out.is_leaf.assign(M, 0);
  out.node_c.assign(3 * M, 0.0);
  out.node_s.assign(M, 0.0);
  for (size_t g = 0; g < G; ++g) {
    const HostTree& t = parts[g];
    for (int64_t i = 0; i < (int64_t)t.level.size(); ++i) {
      const int64_t k = nid[g][i];
      out.level[k] = t.level[i];
      out.begin[k] = t.begin[i] + pt_off[g];
      out.end[k] = t.end[i] + pt_off[g];
      out.parent[k] = t.parent[i] < 0 ? -1 : nid[g][t.parent[i]];
      out.first_child[k] = t.first_child[i] < 0 ? -1 : (int32_t)nid[g][t.first_child[i]];
      out.n_child[k] = t.n_child[i];
      out.is_leaf[k] = t.is_leaf[i];
      out.node_s[k] = t.node_s[i];
      for (int a = 0; a < 3; ++a) {
        out.coord[3 * k + a] = t.coord[3 * i + a];
        out.node_c[3 * k + a] = t.node_c[3 * i + a];
      }
    }
    for (int64_t i = 0; i < t.n; ++i) out.perm.push_back(t.perm[i] + pt_off[g]);
    out.geom_center.insert(out.geom_center.end(), t.center, t.center + 3);
  }
  // lists follow their target node
  std::vector<int64_t> owner_g(M), owner_i(M);
  for (size_t g = 0; g < G; ++g)
    for (int64_t i = 0; i < (int64_t)parts[g].level.size(); ++i) {
      owner_g[nid[g][i]] = (int64_t)g;
      owner_i[nid[g][i]] = i;
    }
  out.v_ptr.assign(M + 1, 0);
  out.u_ptr.assign(M + 1, 0);
  for (int64_t k = 0; k < M; ++k) {
    const HostTree& t = parts[owner_g[k]];
    const auto& map = nid[owner_g[k]];
    const int64_t i = owner_i[k];
    for (int64_t q = t.v_ptr[i]; q < t.v_ptr[i + 1]; ++q) {
      out.v_src.push_back(map[t.v_src[q]]);
      out.v_off.push_back(t.v_off[q]);
    }
    out.v_ptr[k + 1] = (int64_t)out.v_src.size();
    for (int64_t q = t.u_ptr[i]; q < t.u_ptr[i + 1]; ++q) out.u_src.push_back(map[t.u_src[q]]);
    out.u_ptr[k + 1] = (int64_t)out.u_src.size();
  }
  out.center[0] = out.center[1] = out.center[2] = 0.0;
  out.half = 0.0;
  return ED_OK;
}

}  // namespace ed
