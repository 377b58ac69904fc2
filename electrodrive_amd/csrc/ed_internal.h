// Internal declarations shared by the native library's translation units.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ednative.h"

namespace ed {

// Coulomb constant 1/(4 pi eps0), eps0 = 8.854187817e-12, exactly as the
// double the reference computes (electrodrive/utils/config.py:34-35):
// 8987551787.997911.
constexpr double kKE = 0x1.0bd99415ffbb9p+33;

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

// Complex coefficient count (m >= 0 half) and real packed count for order p.
inline int n_complex(int p) { return (p + 1) * (p + 2) / 2; }
inline int n_real(int p) { return (p + 1) * (p + 1); }

// ---------------------------------------------------------------------------
// Host lattice octree (built once per geometry).
// ---------------------------------------------------------------------------
struct HostTree {
  int64_t n = 0;
  int leaf_size = 64, max_level = 16, ws = 1;
  double center[3] = {0, 0, 0};
  double half = 0;                    // root cube half side
  std::vector<int64_t> perm;          // tree position -> original index
  std::vector<uint64_t> keys;         // sorted level-21 Morton keys
  // nodes, stored level by level, Morton-sorted within a level
  std::vector<int32_t> level;
  std::vector<int64_t> begin, end, parent;
  std::vector<int32_t> first_child, n_child;
  std::vector<int32_t> coord;         // 3 per node (integer lattice coords)
  std::vector<uint8_t> is_leaf;
  std::vector<int64_t> level_begin;   // node range of each level (size depth+2)
  int depth = 0;
  // lists
  std::vector<int64_t> v_ptr, v_src;  // CSR over target nodes
  std::vector<int32_t> v_off;         // offset slot of each V entry
  std::vector<int64_t> u_ptr, u_src;  // CSR over all nodes (empty for non-leaves)
  // node geometry in the frame of its own geometry (root centre at 0)
  std::vector<double> node_c;         // 3 per node
  std::vector<double> node_s;         // box side per node
  // geometries (one for a single plan; several for a batched plan), in tree
  // order each geometry's points form one contiguous range
  std::vector<double> geom_center;    // 3 per geometry
  std::vector<int64_t> geom_pt_off;   // n_geoms + 1
  int64_t n_leaves() const;
};

int build_host_tree(const double* centroids, int64_t n, int leaf_size,
                    int max_level, int ws, HostTree& t);
// Concatenate per-geometry trees (point offsets = running sums of part.n)
// into one tree whose nodes are ordered by level, then geometry.
int merge_host_trees(std::vector<HostTree>& parts, HostTree& out);

// Offset slot encoding: o in [-R, R]^3, R = 2 ws + 1.
inline int offset_radius(int ws) { return 2 * ws + 1; }
inline int offset_slot(int ws, int ox, int oy, int oz) {
  const int R = offset_radius(ws), W = 2 * R + 1;
  return (ox + R) + W * ((oy + R) + W * (oz + R));
}

// ---------------------------------------------------------------------------
// Host operator tables (FP64, row-major (p+1)^2 x (p+1)^2).
// ---------------------------------------------------------------------------
void m2l_matrix(int p, double ox, double oy, double oz, double* out);
void m2m_matrix(int p, int octant, double* out);
void l2l_matrix(int p, int octant, double* out);

// Host Morton keys, reference arithmetic (fmm3d/geometry.py:671-799).
void morton_keys_host(const double* pts, int64_t n, int level,
                      const double* bbox_center, const double* bbox_half,
                      int64_t* keys);

}  // namespace ed
