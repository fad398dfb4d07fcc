// extern "C" surface of the host-side library (libchanga_host.so): tree
// build, node geometry, interaction lists and the Ewald h-loop table. These
// are the caller-side steps that precede the GPU work requests.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "host_tree.h"

using namespace cgh;

extern "C" {

// ---- tree ------------------------------------------------------------------
void* cgh_tree_build(int n, const double* pos, const double* box_lo, double box_size,
                     int bucket_size) {
  return build_tree(n, pos, box_lo, box_size, bucket_size);
}

// out: n_nodes, n_buckets, n_depths
void cgh_tree_sizes(void* h, int32_t* out) {
  Tree* t = (Tree*)h;
  out[0] = t->n_nodes;
  out[1] = (int32_t)t->bucket_node.size();
  out[2] = (int32_t)t->level_offsets.size() - 1;
}

void cgh_tree_export(void* h, int32_t* perm, int32_t* first, int32_t* count, int32_t* child,
                     int32_t* depth, double* box_radius, int32_t* level_offsets,
                     int32_t* bucket_node) {
  Tree* t = (Tree*)h;
  std::memcpy(perm, t->perm.data(), sizeof(int32_t) * t->n);
  std::memcpy(first, t->first.data(), sizeof(int32_t) * t->n_nodes);
  std::memcpy(count, t->count.data(), sizeof(int32_t) * t->n_nodes);
  std::memcpy(child, t->child.data(), sizeof(int32_t) * 2 * t->n_nodes);
  std::memcpy(depth, t->depth.data(), sizeof(int32_t) * t->n_nodes);
  std::memcpy(box_radius, t->box_radius.data(), sizeof(double) * t->n_nodes);
  std::memcpy(level_offsets, t->level_offsets.data(), sizeof(int32_t) * t->level_offsets.size());
  std::memcpy(bucket_node, t->bucket_node.data(), sizeof(int32_t) * t->bucket_node.size());
}

void cgh_tree_free(void* h) { delete (Tree*)h; }

void cgh_node_geometry(int n_nodes, const int32_t* first, const int32_t* count,
                       const int32_t* child, const double* box_radius, const double* pos,
                       const double* mass, const double* soft, double* geo) {
  node_geometry(n_nodes, first, count, child, box_radius, pos, mass, soft, geo);
}

// ---- interaction lists -------------------------------------------------------
void* cgh_walk(int n_buckets, const int32_t* bucket_node, const int32_t* node_count,
               const int32_t* node_first, const int32_t* child, const double* geo,
               double theta, int n_replicas, double period, int n_threads) {
  if (n_buckets >= (1 << 22)) return nullptr;  // encodeOffset limit
  WalkParams p;
  p.theta = theta;
  p.n_replicas = n_replicas;
  p.period = period;
  p.n_threads = n_threads;
  return walk(n_buckets, bucket_node, node_count, node_first, child, geo, p);
}

// out: n_cells, n_parts, n_part_sources
void cgh_lists_sizes(void* h, int64_t* out) {
  Lists* L = (Lists*)h;
  out[0] = (int64_t)L->cells.size() / 2;
  out[1] = (int64_t)L->parts.size() / 3;
  out[2] = L->n_part_sources;
}

void cgh_lists_export(void* h, int32_t* cells, int64_t* cell_marks, int64_t* soft_start,
                      int32_t* parts, int64_t* part_marks) {
  Lists* L = (Lists*)h;
  std::memcpy(cells, L->cells.data(), sizeof(int32_t) * L->cells.size());
  std::memcpy(cell_marks, L->cell_marks.data(), sizeof(int64_t) * (L->n_buckets + 1));
  std::memcpy(soft_start, L->soft_start.data(), sizeof(int64_t) * L->n_buckets);
  std::memcpy(parts, L->parts.data(), sizeof(int32_t) * L->parts.size());
  std::memcpy(part_marks, L->part_marks.data(), sizeof(int64_t) * (L->n_buckets + 1));
}

void cgh_lists_free(void* h) { delete (Lists*)h; }

// ---- root moments for the Ewald record ------------------------------------------
// momRescaleFmomr to unit scale followed by momFmomr2Momc (Ewald.cpp:295-296,
// moments.c:1563-1600): the 32 complete components in MOMC order.
void cgh_root_momc(const double* q, double u, double* mc) {
  const double s[5] = {1.0, u, u * u, u * u * u, u * u * u * u};
  const int order[22] = {0, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 3, 3, 4, 4, 4, 4, 4, 4, 4, 4, 4};
  for (int k = 0; k < 22; ++k) mc[k] = q[k] * s[order[k]];
  mc[22] = -(mc[1] + mc[2]);                // zz
  mc[23] = -(mc[6] + mc[7]);                // xzz
  mc[24] = -(mc[8] + mc[9]);                // yzz
  mc[25] = -(mc[10] + mc[11]);              // zzz
  mc[26] = -(mc[13] + mc[19]);              // xxzz
  mc[27] = -(mc[15] + mc[14]);              // xyzz
  mc[28] = -(mc[17] + mc[21]);              // xzzz
  mc[29] = -(mc[19] + mc[16]);              // yyzz
  mc[30] = -(mc[20] + mc[18]);              // yzzz
  mc[31] = -(mc[26] + mc[29]);              // zzzz
}

// ---- Ewald h-loop table --------------------------------------------------------
// TreePiece::EwaldInit (Ewald.cpp:285-383): the root's radius-scaled reduced
// moments are brought to unit scale and completed by tracelessness
// (momRescaleFmomr + momFmomr2Momc); for every integer h with
// 0 < |h|^2 <= hcut^2 (hx outermost, hz innermost) the table holds
// k = 2 pi h / L and the cosine / sine factors
//   hC = -w (M - (Q2:kk)/2 + (Q4:kkkk)/24),  hS = w (Q3:kkk)/6,
//   w = exp(-k^2 / 4 alpha^2) / (pi h^2 L),  alpha = 2 / L,
// which are what QEVAL (Ewald.cpp:6-46) computes from the gam[] weights with
// the odd (resp. even) orders zeroed. ewt[5*i] = kx, ky, kz, hC, hS.
int cgh_ewald_table(const double* q, double u_root, double L, double hcut, double* ewt,
                    int max_n) {
  // complete components at unit scale, MOMC order (moments.h:42-50)
  double s2 = u_root * u_root, s3 = s2 * u_root, s4 = s3 * u_root;
  double m = q[0];
  double xx = q[1] * s2, yy = q[2] * s2, xy = q[3] * s2, xz = q[4] * s2, yz = q[5] * s2;
  double zz = -(xx + yy);
  double xxx = q[6] * s3, xyy = q[7] * s3, xxy = q[8] * s3, yyy = q[9] * s3, xxz = q[10] * s3,
         yyz = q[11] * s3, xyz = q[12] * s3;
  double xzz = -(xxx + xyy), yzz = -(xxy + yyy), zzz = -(xxz + yyz);
  double xxxx = q[13] * s4, xyyy = q[14] * s4, xxxy = q[15] * s4, yyyy = q[16] * s4,
         xxxz = q[17] * s4, yyyz = q[18] * s4, xxyy = q[19] * s4, xxyz = q[20] * s4,
         xyyz = q[21] * s4;
  double xxzz = -(xxxx + xxyy), xyzz = -(xxxy + xyyy), xzzz = -(xxxz + xyyz),
         yyzz = -(xxyy + yyyy), yzzz = -(xxyz + yyyz), zzzz = -(xxzz + yyzz);
  const double PI = 3.14159265358979323846;
  double alpha = 2.0 / L;
  double k4 = PI * PI / (alpha * alpha * L * L);
  int hreps = (int)std::ceil(hcut);
  int n = 0;
  for (int hx = -hreps; hx <= hreps; ++hx)
    for (int hy = -hreps; hy <= hreps; ++hy)
      for (int hz = -hreps; hz <= hreps; ++hz) {
        int h2 = hx * hx + hy * hy + hz * hz;
        if (h2 == 0 || h2 > hcut * hcut) continue;
        if (n >= max_n) return -1;
        double f = 2.0 * PI / L;
        double x = f * hx, y = f * hy, z = f * hz;
        double w = std::exp(-k4 * h2) / (PI * h2 * L);
        // Q2 : kk
        double q2 = xx * x * x + yy * y * y + zz * z * z + 2 * (xy * x * y + xz * x * z + yz * y * z);
        // Q3 : kkk
        double q3 = xxx * x * x * x + yyy * y * y * y + zzz * z * z * z +
                    3 * (xxy * x * x * y + xxz * x * x * z + xyy * x * y * y + yyz * y * y * z +
                         xzz * x * z * z + yzz * y * z * z) +
                    6 * xyz * x * y * z;
        // Q4 : kkkk
        double q4 = xxxx * x * x * x * x + yyyy * y * y * y * y + zzzz * z * z * z * z +
                    4 * (xxxy * x * x * x * y + xxxz * x * x * x * z + xyyy * x * y * y * y +
                         yyyz * y * y * y * z + xzzz * x * z * z * z + yzzz * y * z * z * z) +
                    6 * (xxyy * x * x * y * y + xxzz * x * x * z * z + yyzz * y * y * z * z) +
                    12 * (xxyz * x * x * y * z + xyyz * x * y * y * z + xyzz * x * y * z * z);
        ewt[5 * n + 0] = x;
        ewt[5 * n + 1] = y;
        ewt[5 * n + 2] = z;
        ewt[5 * n + 3] = -w * (m - 0.5 * q2 + q4 / 24.0);
        ewt[5 * n + 4] = w * q3 / 6.0;
        ++n;
      }
  return n;
}

}  // extern "C"
