// Host-side tree, geometry and interaction-list builder (caller side of the
// GPU boundary). See host_tree.cpp and host_walk.cpp.
#pragma once
#include <cstdint>
#include <vector>

namespace cgh {

struct Tree {
  int n = 0, bucket_size = 0;
  int32_t n_nodes = 0;
  double box_lo[3] = {0, 0, 0}, box_size = 1.0;
  std::vector<int32_t> perm;           // tree position -> input index
  std::vector<int32_t> first, count;   // per node
  std::vector<int32_t> child;          // 2 per node, -1 for a bucket
  std::vector<int32_t> depth;          // per node
  std::vector<double> box_radius;      // per node: half diagonal of its Morton cell
  std::vector<int32_t> level_offsets;  // nodes of depth d: [lo[d], lo[d+1])
  std::vector<int32_t> bucket_node;    // bucket b -> node
  std::vector<int32_t> node_bucket;    // node -> bucket or -1
};

Tree* build_tree(int n, const double* pos, const double box_lo[3], double box_size, int bucket_size);

void node_geometry(int n_nodes, const int32_t* first, const int32_t* count, const int32_t* child,
                   const double* box_radius, const double* pos, const double* mass,
                   const double* soft, double* geo);

struct WalkParams {
  double theta = 0.7;
  int n_replicas = 0;
  double period = 1.0;
  int n_threads = 0;
};

struct Lists {
  int32_t n_buckets = 0;
  std::vector<int32_t> cells;       // ILCell pairs {node, offsetID}
  std::vector<int64_t> cell_marks;  // n_buckets + 1
  std::vector<int64_t> soft_start;  // per bucket: first softened cell (absolute index)
  std::vector<int32_t> parts;       // ILPart triples {first, offsetID, num}
  std::vector<int64_t> part_marks;  // n_buckets + 1
  int64_t n_part_sources = 0;       // sum of num over ILPart entries
};

Lists* walk(int n_buckets, const int32_t* bucket_node, const int32_t* node_count,
            const int32_t* node_first, const int32_t* child, const double* geo,
            const WalkParams& p);

}  // namespace cgh
