// Reference-mode octree and interaction lists (host, bit-exact indexing).
//
// The FMM of this library runs on a lattice octree (ed_tree.cpp), where every
// M2L translation depends on an integer offset.  This file builds the
// reference's own adaptive tree and lists so that callers of
// build_fmm_tree / build_interaction_lists get the same node ids, ranges,
// permutation and pair sets:
//
// - tree (fmm3d/tree.py:801-960, SoA :286-339): tight per-node box,
//   centre 0.5 (min + max), half max(0.5 (max - min), 0) in the points'
//   dtype; leaf iff n <= leaf_size, level >= max_depth or max(half) < 1e-7;
//   split by point > centre (strict), octant x + 2y + 4z, stable partition;
//   node ids in depth-first pre-order with children in octant order (an
//   explicit stack replaces the recursion); radius |half| (1 + 1e-12).
// - lists (fmm3d/interaction_lists.py:273-528): dual traversal from
//   (root, root), LIFO; a pair is admissible (M2L) iff it is not a node with
//   itself of one tree, dist > 0 and max(2 r_s, 2 r_t) / dist < theta (double
//   arithmetic on the dtype's norms); leaf-leaf pairs go to P2P; otherwise
//   the side to refine follows :306-337.  Pairs are deduplicated and sorted
//   by (tgt, src), or (src, tgt) without sort_pairs.
//
// Floating-point contraction is off in this file: every sum and product is
// rounded as written, as in the reference's tensor arithmetic.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "ed_internal.h"

#pragma clang fp contract(off)

struct ed_ref_tree {
  int is_f32 = 0;
  int64_t n = 0;
  int leaf_size = 64, max_depth = 32;
  std::vector<int64_t> index_map;
  // node SoA (pre-order)
  std::vector<double> centers, half, radii;  // 3, 3, 1 per node (values exact in the dtype)
  std::vector<int32_t> levels;
  std::vector<int64_t> parents, children, ranges;  // 1, 8, 2 per node
  std::vector<uint8_t> is_leaf;
  int64_t n_leaves = 0;
  int max_level = 0;
};

struct ed_ref_lists {
  std::vector<int64_t> p2p, m2l;  // (src, tgt) pairs, flattened
  int64_t mac_accept = 0, mac_reject = 0, leaf_leaf = 0, stack_pops = 0, max_stack = 0;
  int truncated = 0;
};

namespace ed {
namespace {

template <typename T>
T norm3(T x, T y, T z) {
  const T s = (T)((T)(x * x + y * y) + z * z);
  return (T)std::sqrt(s);
}

template <typename T>
int build_ref_tree(const T* pts, int64_t n, int leaf_size, int max_depth, ed_ref_tree& t) {
  t.n = n;
  t.leaf_size = leaf_size;
  t.max_depth = max_depth;
  t.index_map.resize((size_t)n);
  for (int64_t i = 0; i < n; ++i) t.index_map[(size_t)i] = i;
  if (n == 0) return ED_OK;
  std::vector<int64_t> scratch((size_t)n);
  struct Task {
    int64_t start, end, parent;
    int level, octant;
  };
  std::vector<Task> stack;
  stack.push_back({0, n, -1, 0, -1});
  while (!stack.empty()) {
    const Task k = stack.back();
    stack.pop_back();
    const int64_t me = (int64_t)t.levels.size();
    if (k.parent >= 0) t.children[(size_t)(8 * k.parent + k.octant)] = me;
    T mn[3], mx[3];
    for (int d = 0; d < 3; ++d) {
      mn[d] = mx[d] = pts[3 * t.index_map[(size_t)k.start] + d];
    }
    for (int64_t q = k.start + 1; q < k.end; ++q) {
      const T* p = pts + 3 * t.index_map[(size_t)q];
      for (int d = 0; d < 3; ++d) {
        mn[d] = std::min(mn[d], p[d]);
        mx[d] = std::max(mx[d], p[d]);
      }
    }
    T c[3], h[3];
    for (int d = 0; d < 3; ++d) {
      c[d] = (T)((T)0.5 * (T)(mn[d] + mx[d]));
      const T hh = (T)((T)0.5 * (T)(mx[d] - mn[d]));
      h[d] = hh > (T)0 ? hh : (T)0;
    }
    for (int d = 0; d < 3; ++d) {
      t.centers.push_back((double)c[d]);
      t.half.push_back((double)h[d]);
    }
    t.radii.push_back((double)(T)(norm3<T>(h[0], h[1], h[2]) * (T)(1.0 + 1e-12)));
    t.levels.push_back(k.level);
    t.parents.push_back(k.parent);
    for (int o = 0; o < 8; ++o) t.children.push_back(-1);
    t.ranges.push_back(k.start);
    t.ranges.push_back(k.end);
    t.max_level = std::max(t.max_level, k.level);
    const int64_t np = k.end - k.start;
    const double hmax = (double)std::max(h[0], std::max(h[1], h[2]));
    if (np <= leaf_size || k.level >= max_depth || hmax < 1e-7) {
      t.is_leaf.push_back(1);
      ++t.n_leaves;
      continue;
    }
    t.is_leaf.push_back(0);
    // stable counting partition by octant
    int64_t cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<uint8_t> oct((size_t)np);
    for (int64_t q = k.start; q < k.end; ++q) {
      const T* p = pts + 3 * t.index_map[(size_t)q];
      const int o = (p[0] > c[0] ? 1 : 0) + (p[1] > c[1] ? 2 : 0) + (p[2] > c[2] ? 4 : 0);
      oct[(size_t)(q - k.start)] = (uint8_t)o;
      ++cnt[o];
    }
    int64_t beg[9];
    beg[0] = k.start;
    for (int o = 0; o < 8; ++o) beg[o + 1] = beg[o] + cnt[o];
    int64_t fill[8];
    for (int o = 0; o < 8; ++o) fill[o] = beg[o];
    for (int64_t q = k.start; q < k.end; ++q)
      scratch[(size_t)(fill[oct[(size_t)(q - k.start)]]++)] = t.index_map[(size_t)q];
    std::memcpy(t.index_map.data() + k.start, scratch.data() + k.start, sizeof(int64_t) * (size_t)np);
    // children in octant order, depth first: push in reverse
    for (int o = 7; o >= 0; --o)
      if (beg[o + 1] > beg[o]) stack.push_back({beg[o], beg[o + 1], me, k.level + 1, o});
  }
  return ED_OK;
}

// Norm of the centre difference in the tree's dtype, returned as double.
double center_dist(const ed_ref_tree& s, int64_t i, const ed_ref_tree& t, int64_t j) {
  if (s.is_f32) {
    const float dx = (float)((float)s.centers[3 * i] - (float)t.centers[3 * j]);
    const float dy = (float)((float)s.centers[3 * i + 1] - (float)t.centers[3 * j + 1]);
    const float dz = (float)((float)s.centers[3 * i + 2] - (float)t.centers[3 * j + 2]);
    return (double)norm3<float>(dx, dy, dz);
  }
  const double dx = s.centers[3 * i] - t.centers[3 * j];
  const double dy = s.centers[3 * i + 1] - t.centers[3 * j + 1];
  const double dz = s.centers[3 * i + 2] - t.centers[3 * j + 2];
  return norm3<double>(dx, dy, dz);
}

int build_ref_lists(const ed_ref_tree& src, const ed_ref_tree& tgt, bool same, double theta,
                    int64_t max_pairs, bool sort_pairs, ed_ref_lists& L) {
  const int64_t ms = (int64_t)src.levels.size(), mt = (int64_t)tgt.levels.size();
  L.max_stack = 1;
  if (ms == 0 || mt == 0) return ED_OK;
  std::vector<std::pair<int64_t, int64_t>> p2p, m2l;
  std::vector<std::pair<int64_t, int64_t>> stack;
  stack.emplace_back(0, 0);
  auto diam = [](const ed_ref_tree& tr, int64_t i) { return 2.0 * tr.radii[(size_t)i]; };
  auto push = [&](int64_t s, int64_t t) {
    if (max_pairs >= 0 && (int64_t)(p2p.size() + m2l.size()) >= max_pairs) {
      L.truncated = 1;
      return;
    }
    stack.emplace_back(s, t);
    L.max_stack = std::max<int64_t>(L.max_stack, (int64_t)stack.size());
  };
  while (!stack.empty()) {
    const auto [s, t] = stack.back();
    stack.pop_back();
    ++L.stack_pops;
    const bool ls = src.is_leaf[(size_t)s] != 0, lt = tgt.is_leaf[(size_t)t] != 0;
    bool mac = false;
    if (!(same && s == t)) {
      const double d = center_dist(src, s, tgt, t);
      if (d != 0.0) mac = std::max(diam(src, s), diam(tgt, t)) / d < theta;
      if (mac) ++L.mac_accept; else ++L.mac_reject;
    }
    if (mac) {
      m2l.emplace_back(s, t);
      continue;
    }
    if (ls && lt) {
      ++L.leaf_leaf;
      p2p.emplace_back(s, t);
      continue;
    }
    bool refine_src;
    if (ls && !lt) {
      refine_src = false;
    } else if (lt && !ls) {
      refine_src = true;
    } else {
      const double ds = diam(src, s), dt = diam(tgt, t);
      if (ds > dt * 1.01) {
        refine_src = true;
      } else if (dt > ds * 1.01) {
        refine_src = false;
      } else {
        const int64_t ns = src.ranges[(size_t)(2 * s + 1)] - src.ranges[(size_t)(2 * s)];
        const int64_t nt = tgt.ranges[(size_t)(2 * t + 1)] - tgt.ranges[(size_t)(2 * t)];
        refine_src = ns >= nt;
      }
    }
    if (refine_src) {
      for (int o = 0; o < 8; ++o) {
        const int64_t ch = src.children[(size_t)(8 * s + o)];
        if (ch >= 0) push(ch, t);
      }
    } else {
      for (int o = 0; o < 8; ++o) {
        const int64_t ch = tgt.children[(size_t)(8 * t + o)];
        if (ch >= 0) push(s, ch);
      }
    }
  }
  auto finish = [&](std::vector<std::pair<int64_t, int64_t>>& v, std::vector<int64_t>& out) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    if (sort_pairs)
      std::stable_sort(v.begin(), v.end(), [](const auto& a, const auto& b) {
        return a.second != b.second ? a.second < b.second : a.first < b.first;
      });
    out.resize(2 * v.size());
    for (size_t k = 0; k < v.size(); ++k) {
      out[2 * k] = v[k].first;
      out[2 * k + 1] = v[k].second;
    }
  };
  finish(p2p, L.p2p);
  finish(m2l, L.m2l);
  return ED_OK;
}

}  // namespace
}  // namespace ed

extern "C" int ed_ref_tree_create(const void* points, int32_t is_f32, int64_t n, int32_t leaf_size,
                                  int32_t max_depth, ed_ref_tree** out) {
  using namespace ed;
  if (!out) return fail(ED_ERR_ARG, "ref tree: NULL output");
  *out = nullptr;
  if (n < 0 || (n > 0 && !points)) return fail(ED_ERR_ARG, "ref tree: bad points");
  if (leaf_size <= 0) return fail(ED_ERR_ARG, "leaf_size must be positive");
  if (max_depth <= 0 || max_depth > 64) return fail(ED_ERR_ARG, "max_depth must be in [1, 64]");
  for (int64_t i = 0; i < 3 * n; ++i) {
    const double v = is_f32 ? (double)static_cast<const float*>(points)[i]
                            : static_cast<const double*>(points)[i];
    if (!std::isfinite(v)) return fail(ED_ERR_ARG, "ref tree: points must be finite");
  }
  ed_ref_tree* t = new (std::nothrow) ed_ref_tree();
  if (!t) return fail(ED_ERR_ALLOC, "ref tree: host allocation");
  t->is_f32 = is_f32 ? 1 : 0;
  int rc;
  try {
    rc = is_f32 ? build_ref_tree<float>(static_cast<const float*>(points), n, leaf_size, max_depth, *t)
                : build_ref_tree<double>(static_cast<const double*>(points), n, leaf_size, max_depth, *t);
  } catch (const std::bad_alloc&) {
    rc = fail(ED_ERR_ALLOC, "ref tree: host allocation");
  }
  if (rc != ED_OK) {
    delete t;
    return rc;
  }
  *out = t;
  return ED_OK;
}

extern "C" int ed_ref_tree_sizes(const ed_ref_tree* t, int64_t* out4) {
  if (!t || !out4) return ed::fail(ED_ERR_ARG, "ref tree: NULL argument");
  out4[0] = t->n;
  out4[1] = (int64_t)t->levels.size();
  out4[2] = t->n_leaves;
  out4[3] = t->levels.empty() ? -1 : t->max_level;
  return ED_OK;
}

extern "C" int ed_ref_tree_export(const ed_ref_tree* t, int64_t* index_map, void* centers,
                                  void* half_extents, void* radii, int32_t* levels,
                                  int64_t* parents, int64_t* children, int64_t* ranges,
                                  uint8_t* is_leaf) {
  if (!t) return ed::fail(ED_ERR_ARG, "ref tree: NULL tree");
  const size_t m = t->levels.size();
  auto put = [&](void* dst, const std::vector<double>& src) {
    if (!dst) return;
    if (t->is_f32) {
      float* d = static_cast<float*>(dst);
      for (size_t k = 0; k < src.size(); ++k) d[k] = (float)src[k];
    } else {
      std::memcpy(dst, src.data(), sizeof(double) * src.size());
    }
  };
  if (index_map && t->n) std::memcpy(index_map, t->index_map.data(), sizeof(int64_t) * (size_t)t->n);
  put(centers, t->centers);
  put(half_extents, t->half);
  put(radii, t->radii);
  if (m) {
    if (levels) std::memcpy(levels, t->levels.data(), sizeof(int32_t) * m);
    if (parents) std::memcpy(parents, t->parents.data(), sizeof(int64_t) * m);
    if (children) std::memcpy(children, t->children.data(), sizeof(int64_t) * 8 * m);
    if (ranges) std::memcpy(ranges, t->ranges.data(), sizeof(int64_t) * 2 * m);
    if (is_leaf) std::memcpy(is_leaf, t->is_leaf.data(), m);
  }
  return ED_OK;
}

extern "C" int ed_ref_tree_destroy(ed_ref_tree* t) {
  delete t;
  return ED_OK;
}

extern "C" int ed_ref_lists_create(const ed_ref_tree* src, const ed_ref_tree* tgt, double theta,
                                   int64_t max_pairs, int32_t sort_pairs, ed_ref_lists** out) {
  using namespace ed;
  if (!out || !src) return fail(ED_ERR_ARG, "ref lists: NULL argument");
  *out = nullptr;
  if (!(theta > 0.0)) return fail(ED_ERR_ARG, "mac_theta must be positive");
  if (theta >= 1.0) return fail(ED_ERR_ARG, "mac_theta should be < 1.0 for a meaningful MAC");
  const bool same = tgt == nullptr || tgt == src;
  if (!same && src->is_f32 != tgt->is_f32) return fail(ED_ERR_ARG, "ref lists: tree dtypes differ");
  ed_ref_lists* L = new (std::nothrow) ed_ref_lists();
  if (!L) return fail(ED_ERR_ALLOC, "ref lists: host allocation");
  int rc;
  try {
    rc = build_ref_lists(*src, same ? *src : *tgt, same, theta, max_pairs, sort_pairs != 0, *L);
  } catch (const std::bad_alloc&) {
    rc = fail(ED_ERR_ALLOC, "ref lists: host allocation");
  }
  if (rc != ED_OK) {
    delete L;
    return rc;
  }
  *out = L;
  return ED_OK;
}

extern "C" int ed_ref_lists_sizes(const ed_ref_lists* L, int64_t* out8) {
  if (!L || !out8) return ed::fail(ED_ERR_ARG, "ref lists: NULL argument");
  out8[0] = (int64_t)(L->p2p.size() / 2);
  out8[1] = (int64_t)(L->m2l.size() / 2);
  out8[2] = L->mac_accept;
  out8[3] = L->mac_reject;
  out8[4] = L->leaf_leaf;
  out8[5] = L->stack_pops;
  out8[6] = L->max_stack;
  out8[7] = L->truncated;
  return ED_OK;
}

extern "C" int ed_ref_lists_export(const ed_ref_lists* L, int64_t* p2p_pairs, int64_t* m2l_pairs) {
  if (!L) return ed::fail(ED_ERR_ARG, "ref lists: NULL lists");
  if (p2p_pairs && !L->p2p.empty())
    std::memcpy(p2p_pairs, L->p2p.data(), sizeof(int64_t) * L->p2p.size());
  if (m2l_pairs && !L->m2l.empty())
    std::memcpy(m2l_pairs, L->m2l.data(), sizeof(int64_t) * L->m2l.size());
  return ED_OK;
}

extern "C" int ed_ref_lists_destroy(ed_ref_lists* L) {
  delete L;
  return ED_OK;
}
