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

// Copies and validates a dfn_problem; computes geometry, loading projections,
// fake-element defaults and fault crossings.  Returns a DFN_* status.
int setup_problem(const dfn_problem *p, HostProblem &hp);

// Builds the H-matrix operator op (DFN_OP_T / DFN_OP_N) of the problem
// (cluster tree, block cluster tree, full-pivot ACA, QR+SVD recompression).
int build_operator(const HostProblem &hp, int op, HostOperator &out);

}  // namespace dfn
