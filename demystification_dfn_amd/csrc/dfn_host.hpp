// Host-side data of one problem: geometry in the reference's element order,
// per-element parameters, fluid-diffusion coefficients and the stress-transfer
// operators as flat block lists.  Built once per handle, then uploaded.
#pragma once
#include <cstdint>
#include <vector>

#include "../../include/dfn.h"

namespace dfn {

// One H-matrix operator as flat block lists in H order (0-based).  Dense block b
// covers rows d_r0..+d_m, cols d_c0..+d_n (column-major data at d_off); low-rank
// block b = U V^T, U (m x k) at l_uoff, V (n x k) at l_voff, column-major.
struct HostOperator {
  int n = 0;
  int nbclus = 0, nbblclus = 0;
  std::vector<int> perm;  // H position -> original element
  std::vector<int> d_r0, d_c0, d_m, d_n;
  std::vector<int64_t> d_off;
  std::vector<double> d_data;
  std::vector<int> l_r0, l_c0, l_m, l_n, l_k;
  std::vector<int64_t> l_uoff, l_voff;
  std::vector<double> l_u, l_v;
  std::vector<int> leaf_r0, leaf_m;  // leaf row clusters (partition of 0..n-1)
  bool valid() const { return n > 0; }
};

struct HostProblem {
  int mode = 3, kernel = 0, nfault = 0, N = 0;
  std::vector<int64_t> fnel, map1;       // per fault: element count, first element (1-based)
  std::vector<double> node;              // 2*(N+1)
  std::vector<double> elem, ds, tang, nrm;  // element midpoints, lengths, tangents, normals
  std::vector<double> a, b, Dc, f0, V0, nld, sld;
  std::vector<double> perm0, perm_star, kmin, Snk;  // permeability data
  std::vector<double> Y0, P0;
  double mu = 0, cs = 0, cp = 0, por = 0, fcomp = 0, rcomp = 0, visc = 0, tol = 0;
  int coupling = 0;
  std::vector<int64_t> src_elem;         // 1-based global element index of each source
  std::vector<double> src_Q, src_tb, src_te;
  std::vector<int64_t> cross;            // ncross*6, crossing(c,1:6)
  std::vector<int64_t> fake;             // 0-based fake element indices
  std::vector<uint8_t> is_fake;
  double eta = 3, eps_aca = 1e-8;
  int nleaf = 100, max_rank = 20;
  int ncross() const { return (int)(cross.size() / 6); }
};

// Geometry in H order for the matrix-entry functions of
// src/hmatrix/eval_kernel_hmatrix.f90 (receiver k, source l): element centres,
// left/right nodes, tangents and normals.
struct HGeom {
  std::vector<double> ex, ey, nlx, nly, nrx, nry, tx, ty, nx, ny;
};

// The partition of one operator (cluster tree, block cluster tree) with its
// leaves in preorder and the geometry in H order; the matrix data are filled
// separately (on the host by build_operator, on the device by
// build_operator_device).
struct HostPartition {
  int n = 0, nbclus = 0, nbblclus = 0;
  bool mode3 = false, normal = false;
  std::vector<int> perm, leaf_r0, leaf_m;
  std::vector<int> d_r0, d_c0, d_m, d_n;  // dense leaves
  std::vector<int> l_r0, l_c0, l_m, l_n;  // admissible leaves
  HGeom g;
};

#define DFN_TRY_HOST(x)          \
  do {                           \
    const int s_ = (x);          \
    if (s_ != DFN_OK) return s_; \
  } while (0)

// Copies and validates a dfn_problem; computes geometry, loading projections,
// fake-element defaults and fault crossings.  Returns a DFN_* status.
int setup_problem(const dfn_problem *p, HostProblem &hp);

// Builds the H-matrix operator op (DFN_OP_T / DFN_OP_N) of the problem
// (cluster tree, block cluster tree, full-pivot ACA, QR+SVD recompression).
int build_operator(const HostProblem &hp, int op, HostOperator &out);

// Pieces of build_operator shared with the device construction.
int build_partition(const HostProblem &hp, int op, HostPartition &P);
// Copies the partition's lists into out and sizes the dense data.
void init_operator(const HostPartition &P, HostOperator &out);
// QR + SVD recompression of one low-rank block (A m x k, B n x k, column-major,
// replaced by the recompressed factors); returns the new rank.
int recompress_block(int m, int n, int k, std::vector<double> &A, std::vector<double> &B, double eps);
// Recompresses nl blocks whose factors arrive packed (block i: A at hA + doa[i],
// m[i] x rank[i]; B at hB + dob[i], n[i] x rank[i]); rank is updated, the
// recompressed factors land in Af / Bf.  OpenMP over blocks.
void recompress_all(const double *hA, const int64_t *doa, const double *hB, const int64_t *dob, const int *m,
                    const int *n, int nl, double eps, std::vector<int> &rank, std::vector<std::vector<double>> &Af,
                    std::vector<std::vector<double>> &Bf);
// Concatenates per-block factors (Af[i]: m x rank[i], Bf[i]: n x rank[i]) into out.
void pack_lowrank(std::vector<std::vector<double>> &Af, std::vector<std::vector<double>> &Bf,
                  const std::vector<int> &rank, HostOperator &out);
// Device construction of the same operator: partition on the host, dense leaves
// and full-pivot ACA on the GPU, recompression on the host.
int build_operator_device(const HostProblem &hp, int op, HostOperator &out, void *stream);

}  // namespace dfn
