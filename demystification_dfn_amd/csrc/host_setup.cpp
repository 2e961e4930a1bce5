// Host-side setup of a problem: copy/validation of the C-ABI description,
// geometry (src/initialisation.f90:139-206), fault crossings (:279-344) and the
// H-matrix construction (src/hmatrix/*.f90) as flat block lists.  Arithmetic on
// values that reach the device keeps the reference's operation order; the file
// is compiled without floating-point contraction.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "dfn_host.hpp"

namespace dfn {

namespace {

template <class T>
std::vector<T> copy_or(const T *src, size_t n, T fill) {
  std::vector<T> v(n, fill);
  if (src) std::copy(src, src + n, v.begin());
  return v;
}

}  // namespace

int setup_problem(const dfn_problem *p, HostProblem &hp) {
  if (!p || p->nb_fault < 1 || p->nb_element < 2 || !p->node || !p->fault_nb_element) return DFN_ERR_ARG;
  if (p->fracture_mode != DFN_MODE_II && p->fracture_mode != DFN_MODE_III) return DFN_ERR_ARG;
  if (p->static_kernel != DFN_KERNEL_DIRECT && p->static_kernel != DFN_KERNEL_HMATRIX) return DFN_ERR_ARG;
  int64_t total = -1;
  for (int f = 0; f < p->nb_fault; f++) {
    if (p->fault_nb_element[f] < 1) return DFN_ERR_ARG;
    total += p->fault_nb_element[f] + 1;
  }
  if (total != p->nb_element) return DFN_ERR_ARG;
  const int N = p->nb_element;
  hp.mode = p->fracture_mode;
  hp.kernel = p->static_kernel;
  hp.nfault = p->nb_fault;
  hp.N = N;
  hp.fnel.assign(p->fault_nb_element, p->fault_nb_element + p->nb_fault);
  hp.map1.resize(hp.nfault);
  int64_t cnt = 0;
  for (int f = 0; f < hp.nfault; f++) {  // src/configuration.f90:237-245
    hp.map1[f] = cnt + 1;
    cnt += hp.fnel[f] + 1;
  }
  hp.node.assign(p->node, p->node + 2 * (size_t)(N + 1));
  hp.a = copy_or(p->a, N, 0.0); hp.b = copy_or(p->b, N, 0.0); hp.Dc = copy_or(p->Dc, N, 0.0);
  hp.f0 = copy_or(p->f0, N, 0.0); hp.V0 = copy_or(p->V0, N, 0.0);
  hp.nld = copy_or(p->normal_loading_dot, N, 0.0);
  hp.sld = copy_or(p->shear_loading_dot, N, 0.0);
  hp.perm0 = copy_or(p->permeability, N, 0.0);
  hp.kmin = copy_or(p->kmin, N, 0.0);
  hp.Snk = copy_or(p->Snk, N, 0.0);
  hp.Y0.assign(3 * (size_t)N, 0.0);
  if (p->V) std::copy(p->V, p->V + N, hp.Y0.begin());
  if (p->theta) std::copy(p->theta, p->theta + N, hp.Y0.begin() + N);
  if (p->sigmaN) std::copy(p->sigmaN, p->sigmaN + N, hp.Y0.begin() + 2 * N);
  hp.P0 = copy_or(p->P, N, 0.0);
  hp.mu = p->mu; hp.cs = p->cs; hp.cp = p->cp;
  hp.por = p->porosity; hp.fcomp = p->fluid_comp; hp.rcomp = p->rock_comp; hp.visc = p->dyn_viscosity;
  hp.tol = p->tol_solver;
  hp.coupling = p->permeability_coupling;
  if (hp.mode == DFN_MODE_II && !(hp.cp > 0)) return DFN_ERR_ARG;
  for (int s = 0; s < p->nb_source; s++) {
    int64_t f = p->source_fault[s], i = p->source_index[s];
    if (f < 1 || f > hp.nfault || i < 1 || i > hp.fnel[f - 1]) return DFN_ERR_ARG;
    hp.src_elem.push_back(hp.map1[f - 1] - 1 + i);  // mapping_fault(f,1)-1+index_injection
    hp.src_Q.push_back(p->source_Q[s]);
    hp.src_tb.push_back(p->source_t_beg[s]);
    hp.src_te.push_back(p->source_t_end[s]);
  }
  hp.eta = p->eta > 0 ? p->eta : 3.0;
  hp.eps_aca = p->eps_aca > 0 ? p->eps_aca : 1e-8;
  hp.nleaf = p->nleaf > 0 ? p->nleaf : 100;
  hp.max_rank = p->max_rank > 0 ? p->max_rank : 20;
  if (hp.max_rank > 64) return DFN_ERR_UNSUPPORTED;

  // fake elements between consecutive faults (RateStateAgingModeIII.f90:105-113)
  hp.is_fake.assign(N, 0);
  int64_t idx = 0;
  for (int f = 0; f < hp.nfault; f++) {
    idx += hp.fnel[f] + 1;
    if (idx <= N) {
      hp.fake.push_back(idx - 1);
      hp.is_fake[idx - 1] = 1;
    }
  }
  for (int64_t i : hp.fake) {  // src/configuration.f90:199-203
    hp.a[i] = hp.b[i] = hp.Dc[i] = hp.f0[i] = hp.V0[i] = 1.0;
    hp.Snk[i] = 1.0;
    hp.nld[i] = hp.sld[i] = hp.perm0[i] = hp.kmin[i] = hp.P0[i] = 0.0;
    hp.Y0[i] = hp.Y0[N + i] = hp.Y0[2 * N + i] = 0.0;
  }

  // geometry, src/initialisation.f90:145-153
  hp.elem.resize(2 * (size_t)N); hp.ds.resize(N); hp.tang.resize(2 * (size_t)N); hp.nrm.resize(2 * (size_t)N);
  for (int k = 0; k < N; k++) {
    const double *n0 = &hp.node[2 * k], *n1 = &hp.node[2 * (k + 1)];
    hp.elem[2 * k] = (n1[0] + n0[0]) * 0.5;
    hp.elem[2 * k + 1] = (n1[1] + n0[1]) * 0.5;
    const double dx = n1[0] - n0[0], dy = n1[1] - n0[1];
    hp.ds[k] = std::sqrt(dx * dx + dy * dy);
    hp.tang[2 * k] = (n1[0] - n0[0]) / hp.ds[k];
    hp.tang[2 * k + 1] = (n1[1] - n0[1]) / hp.ds[k];
    hp.nrm[2 * k] = -hp.tang[2 * k + 1];
    hp.nrm[2 * k + 1] = hp.tang[2 * k];
    if (!(hp.ds[k] > 0)) return DFN_ERR_ARG;
  }
  // loading projections, src/initialisation.f90:159-193
  const double s11 = p->Sigma11_dot, s12 = p->Sigma12_dot, s22 = p->Sigma22_dot;
  const double s31 = p->Sigma31_dot, s32 = p->Sigma32_dot;
  for (int k = 0; k < N; k++) {
    const double t1 = hp.tang[2 * k], t2 = hp.tang[2 * k + 1], n1 = hp.nrm[2 * k], n2 = hp.nrm[2 * k + 1];
    double &sig = hp.Y0[2 * N + k];
    if (hp.mode == DFN_MODE_III) {
      hp.sld[k] = hp.sld[k] - t2 * s31 + t1 * s32;
      sig = sig + n1 * (n1 * s11 + n2 * s12) + n2 * (n1 * s12 + n2 * s22);
    } else {
      sig = sig + n1 * (n1 * 0.0 + n2 * 0.0) + n2 * (n1 * 0.0 + n2 * 0.0);
      hp.nld[k] = hp.nld[k] + n1 * (n1 * s11 + n2 * s12) + n2 * (n1 * s12 + n2 * s22);
      hp.sld[k] = hp.sld[k] + t1 * (n1 * s11 + n2 * s12) + t2 * (n1 * s12 + n2 * s22);
    }
  }
  // permeability_star, src/initialisation.f90:257-261
  hp.perm_star.assign(N, 0.0);
  if (hp.coupling == 1)
    for (int k = 0; k < N; k++)
      hp.perm_star[k] = hp.kmin[k] + (hp.perm0[k] - hp.kmin[k]) *
                                         std::exp(std::fabs(hp.Y0[2 * N + k] + hp.P0[k]) / std::fabs(hp.Snk[k]));

  // fault crossings, src/initialisation.f90:279-344: nodes of two faults closer than
  // 0.01 (single precision) x the smaller of the two faults' minimum element sizes.
  // Fault pairs whose node boxes are farther apart than that bound are skipped.
  std::vector<double> minds(hp.nfault), box(4 * (size_t)hp.nfault);
  for (int f = 0; f < hp.nfault; f++) {
    double mn = INFINITY, x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
    for (int64_t e = 0; e < hp.fnel[f]; e++) mn = std::fmin(mn, hp.ds[hp.map1[f] - 1 + e]);
    for (int64_t q = 0; q <= hp.fnel[f]; q++) {
      const double *pt = &hp.node[2 * (hp.map1[f] - 1 + q)];
      x0 = std::fmin(x0, pt[0]); x1 = std::fmax(x1, pt[0]); y0 = std::fmin(y0, pt[1]); y1 = std::fmax(y1, pt[1]);
    }
    minds[f] = mn;
    box[4 * f] = x0; box[4 * f + 1] = x1; box[4 * f + 2] = y0; box[4 * f + 3] = y1;
  }
  for (int f1 = 0; f1 < hp.nfault; f1++)
    for (int f2 = f1 + 1; f2 < hp.nfault; f2++) {
      const double thr = (double)0.01f * std::fmin(minds[f1], minds[f2]);
      const double gx = std::fmax(box[4 * f1] - box[4 * f2 + 1], box[4 * f2] - box[4 * f1 + 1]);
      const double gy = std::fmax(box[4 * f1 + 2] - box[4 * f2 + 3], box[4 * f2 + 2] - box[4 * f1 + 3]);
      if (gx > thr || gy > thr) continue;
      for (int64_t q1 = 1; q1 <= hp.fnel[f1] + 1; q1++) {
        const double *a1 = &hp.node[2 * (hp.map1[f1] + q1 - 2)];
        for (int64_t q2 = 1; q2 <= hp.fnel[f2] + 1; q2++) {
          const double *a2 = &hp.node[2 * (hp.map1[f2] + q2 - 2)];
          const double d = std::sqrt((a1[0] - a2[0]) * (a1[0] - a2[0]) + (a1[1] - a2[1]) * (a1[1] - a2[1]));
          if (d < thr) {
            const int64_t row[6] = {f1 + 1, f2 + 1, q1 + hp.map1[f1] - 1, q2 + hp.map1[f2] - 1, q1, q2};
            hp.cross.insert(hp.cross.end(), row, row + 6);
          }
        }
      }
    }
  for (int c = 0; c < hp.ncross(); c++)  // the crossing stencil reads elements node-1 and node
    for (int q = 2; q < 4; q++)
      if (hp.cross[6 * c + q] - 1 < 1 || hp.cross[6 * c + q] > N) return DFN_ERR_ARG;
  return DFN_OK;
}

// ---------------------------------------------------------------------------
// H-matrix construction
// ---------------------------------------------------------------------------
namespace {


inline double entry_III(const HGeom &g, int k, int l) {  // eval_kernel_ModeIII :7-52
  const double e1 = g.ex[k], e2 = g.ey[k];
  const double a1 = g.nlx[l] - e1, a2 = g.nly[l] - e2, b1 = g.nrx[l] - e1, b2 = g.nry[l] - e2;
  const double r1 = std::sqrt(a1 * a1 + a2 * a2), r2 = std::sqrt(b1 * b1 + b2 * b2);
  const double I1 = g.tx[k] / (r1 * r1) * (e1 - g.nlx[l]) + g.ty[k] / (r1 * r1) * (e2 - g.nly[l]);
  const double I2 = g.tx[k] / (r2 * r2) * (e1 - g.nrx[l]) + g.ty[k] / (r2 * r2) * (e2 - g.nry[l]);
  return -I1 + I2;
}

inline double entry_II(const HGeom &g, int k, int l, bool normal) {  // eval_kernel_ModeII_T/_N :57-199
  const double e1 = g.ex[k], e2 = g.ey[k];
  const double exl = e1 - g.nlx[l], eyl = e2 - g.nly[l], exr = e1 - g.nrx[l], eyr = e2 - g.nry[l];
  const double rl = std::sqrt(exl * exl + eyl * eyl), rr = std::sqrt(exr * exr + eyr * eyr);
  const double g1l = exl / rl, g2l = eyl / rl, g1r = exr / rr, g2r = eyr / rr;
  const double n1 = g.nx[k], n2 = g.ny[k];
  // normal_left = normal_right = normal (src/initialisation.f90:199-200)
  const double ml1 = g.nx[l], ml2 = g.ny[l];
  if (!normal) {
    const double cl = 4.0 * n1 * n2 * g1l * g2l + (n2 * n2 - n1 * n1) * (g2l * g2l - g1l * g1l);
    const double cr = 4.0 * n1 * n2 * g1r * g2r + (n2 * n2 - n1 * n1) * (g2r * g2r - g1r * g1r);
    const double I1tl = cl * g1l / rl, I2tl = cl * g2l / rl, I1tr = cr * g1r / rr, I2tr = cr * g2r / rr;
    return ml2 * I1tl - ml2 * I1tr - ml1 * I2tl + ml1 * I2tr;
  }
  const double dl = 2.0 * n1 * n2 * (g2l * g2l - g1l * g1l) - (n2 * n2 - n1 * n1) * 2.0 * g1l * g2l;
  const double dr = 2.0 * n1 * n2 * (g2r * g2r - g1r * g1r) - (n2 * n2 - n1 * n1) * 2.0 * g1r * g2r;
  const double I1nl = g1l / rl - g2l / rl * dl, I2nl = g2l / rl + g1l / rl * dl;
  const double I1nr = g1r / rr - g2r / rr * dr, I2nr = g2r / rr + g1r / rr * dr;
  return ml1 * I1nl - ml1 * I1nr + ml2 * I2nl - ml2 * I2nr;
}

struct Cluster {
  int first, last;  // 0-based inclusive
  int son1 = -1, son2 = -1;
  double x0, x1, y0, y1;  // element-centre bounding box
  int size() const { return last - first + 1; }
};

struct TreeBuilder {
  std::vector<double> cx, cy;  // element centres, permuted in place
  std::vector<int> perm;
  std::vector<Cluster> cl;
  int nleaf;
  bool fail = false;

  // One recursion that does what parameters_cluster_tree (reorder) and
  // create_cluster_tree (preorder numbering) do in two passes
  // (src/hmatrix/cluster_tree.f90:69-257, affiliation :400-534, reorder :595-712).
  int make(int f, int l) {
    const int id = (int)cl.size();
    cl.push_back(Cluster{f, l});
    double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
    for (int i = f; i <= l; i++) {
      x0 = std::fmin(x0, cx[i]); x1 = std::fmax(x1, cx[i]); y0 = std::fmin(y0, cy[i]); y1 = std::fmax(y1, cy[i]);
    }
    cl[id].x0 = x0; cl[id].x1 = x1; cl[id].y0 = y0; cl[id].y1 = y1;
    if (l - f + 1 <= nleaf || fail) return id;
    const bool ydir = (y1 - y0) >= (x1 - x0);  // ties go to the y axis
    const double lo = ydir ? y0 : x0, hi = ydir ? y1 : x1;
    const double mid = lo + (hi - lo) / 2;
    std::vector<int> upper, lower;  // coordinate >= mid, < mid, each in index order
    for (int i = f; i <= l; i++) ((ydir ? cy[i] : cx[i]) >= mid ? upper : lower).push_back(i);
    if (upper.empty() || lower.empty()) { fail = true; return id; }
    const std::vector<int> &first = (upper[0] <= lower[0]) ? upper : lower;
    const std::vector<int> &second = (upper[0] <= lower[0]) ? lower : upper;
    std::vector<double> nx, ny;
    std::vector<int> np;
    nx.reserve(l - f + 1); ny.reserve(l - f + 1); np.reserve(l - f + 1);
    for (const auto *grp : {&first, &second})
      for (int i : *grp) { nx.push_back(cx[i]); ny.push_back(cy[i]); np.push_back(perm[i]); }
    std::copy(nx.begin(), nx.end(), cx.begin() + f);
    std::copy(ny.begin(), ny.end(), cy.begin() + f);
    std::copy(np.begin(), np.end(), perm.begin() + f);
    const int split = f + (int)first.size();
    const int s1 = make(f, split - 1);
    const int s2 = make(split, l);
    cl[id].son1 = s1;
    cl[id].son2 = s2;
    return id;
  }
};

struct Block {
  int type;  // 0 dense, 1 low-rank, 2 split
  int row, col;
};

struct BlockTreeBuilder {
  const std::vector<Cluster> &cl;
  double eta;
  std::vector<Block> blocks;
  explicit BlockTreeBuilder(const std::vector<Cluster> &c, double e) : cl(c), eta(e) {}

  // admissibility_condition, src/hmatrix/block_cluster_tree.f90:281-476
  bool admissible(const Cluster &r, const Cluster &c) const {
    double dr = 0.0, dc = 0.0, dist = 0.0;
    dr = dr + (r.x1 - r.x0) * (r.x1 - r.x0);
    dr = dr + (r.y1 - r.y0) * (r.y1 - r.y0);
    dr = std::sqrt(dr);
    dc = dc + (c.x1 - c.x0) * (c.x1 - c.x0);
    dc = dc + (c.y1 - c.y0) * (c.y1 - c.y0);
    dc = std::sqrt(dc);
    double a = std::fmax(0.0, r.x0 - c.x1), b = std::fmax(0.0, c.x0 - r.x1);
    dist = dist + a * a + b * b;
    a = std::fmax(0.0, r.y0 - c.y1); b = std::fmax(0.0, c.y0 - r.y1);
    dist = dist + a * a + b * b;
    dist = std::sqrt(dist);
    return eta * dist >= std::fmax(dr, dc);
  }

  // create_blcluster_tree (:57-183): preorder, sons (r1c1, r1c2, r2c1, r2c2)
  void make(int r, int c) {
    const int id = (int)blocks.size();
    blocks.push_back(Block{0, r, c});
    if (admissible(cl[r], cl[c])) { blocks[id].type = 1; return; }
    if (cl[r].son1 >= 0 && cl[c].son1 >= 0) {
      blocks[id].type = 2;
      for (int rs : {cl[r].son1, cl[r].son2})
        for (int cs : {cl[c].son1, cl[c].son2}) make(rs, cs);
    }
  }
};

// compute_full_ACA (src/hmatrix/ACA_algorithm.f90:665-826) on an assembled block R
// (m x n, column-major, overwritten).  A (m x max_rank), B (n x max_rank) zeroed.
int full_pivot_aca(int m, int n, std::vector<double> &R, double eps, int max_rank, double *A, double *B) {
  const size_t mn = (size_t)m * n;
  double nA = 0.0;
  for (size_t t = 0; t < mn; t++) nA = nA + R[t] * R[t];
  nA = std::sqrt(nA);
  const double piv_tol = (double)1.0e-14f;
  int low_rank = 0, r;
  const int rmax = std::min(m, n);
  // pivot of the current residual: first maximum of |R| in column-major order
  double best = -1.0;
  size_t p = 0;
  for (size_t t = 0; t < mn; t++)
    if (std::fabs(R[t]) > best) { best = std::fabs(R[t]); p = t; }
  for (r = 1; r <= rmax; r++) {
    const int ip = (int)(p % (size_t)m), jp = (int)(p / (size_t)m);
    if (std::fabs(R[p]) <= piv_tol) {
      for (int j = 0; j < n; j++) B[j + (size_t)(r - 1) * n] = 0.0;
      low_rank = r - 1;
      break;
    }
    const double piv = R[p];
    double *a = A + (size_t)(r - 1) * m, *b = B + (size_t)(r - 1) * n;
    for (int i = 0; i < m; i++) a[i] = R[i + (size_t)jp * m];
    for (int j = 0; j < n; j++) b[j] = R[ip + (size_t)j * m] / piv;
    // rank-1 update fused with the residual norm and the next pivot search
    double nR = 0.0;
    best = -1.0;
    p = 0;
    for (int j = 0; j < n; j++) {
      const double bj = b[j];
      double *col = &R[(size_t)j * m];
      for (int i = 0; i < m; i++) {
        const double v = col[i] - a[i] * bj;
        col[i] = v;
        nR = nR + v * v;
        if (std::fabs(v) > best) { best = std::fabs(v); p = (size_t)j * m + i; }
      }
    }
    nR = std::sqrt(nR);
    if (nR <= eps * nA) { low_rank = r; break; }
    if (r + 1 > max_rank) { low_rank = r; break; }
  }
  if (low_rank == 0) low_rank = std::min(r, rmax);
  return low_rank;
}

// Thin Householder QR of M (m x k, column-major, m > k): Q (m x k), R (k x k).
void householder_qr(int m, int k, const double *M, double *Q, double *R) {
  std::vector<double> W(M, M + (size_t)m * k), V((size_t)m * k, 0.0), beta(k, 0.0);
  for (int j = 0; j < k; j++) {
    double s = 0.0;
    for (int i = j; i < m; i++) s += W[i + (size_t)j * m] * W[i + (size_t)j * m];
    const double nrm = std::sqrt(s);
    if (nrm == 0.0) continue;
    double *v = &V[(size_t)j * m];
    const double alpha = W[j + (size_t)j * m] >= 0 ? -nrm : nrm;
    for (int i = j; i < m; i++) v[i] = W[i + (size_t)j * m];
    v[j] -= alpha;
    double vv = 0.0;
    for (int i = j; i < m; i++) vv += v[i] * v[i];
    beta[j] = vv > 0 ? 2.0 / vv : 0.0;
    for (int c = j; c < k; c++) {
      double d = 0.0;
      for (int i = j; i < m; i++) d += v[i] * W[i + (size_t)c * m];
      d *= beta[j];
      for (int i = j; i < m; i++) W[i + (size_t)c * m] -= d * v[i];
    }
  }
  for (int c = 0; c < k; c++)
    for (int r = 0; r < k; r++) R[r + (size_t)c * k] = r <= c ? W[r + (size_t)c * m] : 0.0;
  std::fill(Q, Q + (size_t)m * k, 0.0);
  for (int c = 0; c < k; c++) Q[c + (size_t)c * m] = 1.0;
  for (int j = k - 1; j >= 0; j--) {
    if (beta[j] == 0.0) continue;
    const double *v = &V[(size_t)j * m];
    for (int c = 0; c < k; c++) {
      double d = 0.0;
      for (int i = j; i < m; i++) d += v[i] * Q[i + (size_t)c * m];
      d *= beta[j];
      for (int i = j; i < m; i++) Q[i + (size_t)c * m] -= d * v[i];
    }
  }
}

// One-sided Jacobi SVD of a k x k matrix: A = U diag(S) V^T, S descending.
void jacobi_svd(int k, const double *A, double *U, double *S, double *V) {
  std::vector<double> W(A, A + (size_t)k * k);
  std::fill(V, V + (size_t)k * k, 0.0);
  for (int i = 0; i < k; i++) V[i + (size_t)i * k] = 1.0;
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = 0.0;
    for (int p = 0; p < k - 1; p++)
      for (int q = p + 1; q < k; q++) {
        double al = 0, be = 0, ga = 0;
        for (int i = 0; i < k; i++) {
          const double wp = W[i + (size_t)p * k], wq = W[i + (size_t)q * k];
          al += wp * wp; be += wq * wq; ga += wp * wq;
        }
        if (ga == 0.0) continue;
        const double c0 = std::fabs(ga) / std::sqrt(al * be);
        if (!(c0 > 1e-15)) continue;
        off = std::fmax(off, c0);
        const double zeta = (be - al) / (2.0 * ga);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
        for (int i = 0; i < k; i++) {
          const double wp = W[i + (size_t)p * k], wq = W[i + (size_t)q * k];
          W[i + (size_t)p * k] = c * wp - s * wq;
          W[i + (size_t)q * k] = s * wp + c * wq;
          const double vp = V[i + (size_t)p * k], vq = V[i + (size_t)q * k];
          V[i + (size_t)p * k] = c * vp - s * vq;
          V[i + (size_t)q * k] = s * vp + c * vq;
        }
      }
    if (off <= 1e-15) break;
  }
  std::vector<double> sv(k);
  std::vector<int> ord(k);
  for (int j = 0; j < k; j++) {
    double s = 0.0;
    for (int i = 0; i < k; i++) s += W[i + (size_t)j * k] * W[i + (size_t)j * k];
    sv[j] = std::sqrt(s);
    ord[j] = j;
  }
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return sv[x] > sv[y]; });
  std::vector<double> V2((size_t)k * k);
  for (int j = 0; j < k; j++) {
    const int o = ord[j];
    S[j] = sv[o];
    for (int i = 0; i < k; i++) {
      U[i + (size_t)j * k] = sv[o] > 0 ? W[i + (size_t)o * k] / sv[o] : (i == j ? 1.0 : 0.0);
      V2[i + (size_t)j * k] = V[i + (size_t)o * k];
    }
  }
  std::copy(V2.begin(), V2.end(), V);
}

// compression_low_rank (src/hmatrix/H_matrix_recompression.f90:145-301).
// A (m x k), B (n x k) in, recompressed factors out; returns the new rank.
int recompress(int m, int n, int k, std::vector<double> &A, std::vector<double> &B, double eps) {
  const int mn = std::min(m, n);
  if (k < mn) {
    std::vector<double> Qa((size_t)m * k), Ra((size_t)k * k), Qb((size_t)n * k), Rb((size_t)k * k);
    householder_qr(m, k, A.data(), Qa.data(), Ra.data());
    householder_qr(n, k, B.data(), Qb.data(), Rb.data());
    std::vector<double> R((size_t)k * k, 0.0), U((size_t)k * k), S(k), V((size_t)k * k);
    for (int c = 0; c < k; c++)
      for (int r = 0; r < k; r++) {
        double s = 0.0;
        for (int l = std::max(r, c); l < k; l++) s += Ra[r + (size_t)l * k] * Rb[c + (size_t)l * k];
        R[r + (size_t)c * k] = s;
      }
    jacobi_svd(k, R.data(), U.data(), S.data(), V.data());
    int out = 0;
    for (int i = 0; i < k; i++) {  // keep through the first sigma_i <= eps*sigma_1 (:562-589)
      out++;
      if (S[i] <= eps * S[0]) break;
    }
    std::vector<double> An((size_t)m * out, 0.0), Bn((size_t)n * out, 0.0);
    for (int j = 0; j < out; j++)
      for (int l = 0; l < k; l++) {
        const double ta = U[l + (size_t)j * k] * S[j], tb = V[l + (size_t)j * k];
        for (int i = 0; i < m; i++) An[i + (size_t)j * m] += ta * Qa[i + (size_t)l * m];
        for (int i = 0; i < n; i++) Bn[i + (size_t)j * n] += tb * Qb[i + (size_t)l * n];
      }
    A.swap(An);
    B.swap(Bn);
    return out;
  }
  // rank >= min(m,n): ACA_full_pivot on F = A B^T (:274-297, :305-423)
  std::vector<double> F((size_t)m * n, 0.0);
  for (int j = 0; j < n; j++)
    for (int l = 0; l < k; l++) {
      const double t = B[j + (size_t)l * n];
      for (int i = 0; i < m; i++) F[i + (size_t)j * m] = F[i + (size_t)j * m] + t * A[i + (size_t)l * m];
    }
  std::vector<double> An((size_t)m * mn, 0.0), Bn((size_t)n * mn, 0.0);
  double nF = 0.0;
  for (double v : F) nF = nF + v * v;
  nF = std::sqrt(nF);
  int rank;
  for (rank = 1; rank <= mn; rank++) {
    double mx = 0.0;
    size_t ind = 0;
    for (size_t t = 0; t < F.size(); t++)
      if (std::fabs(F[t]) > mx) { mx = std::fabs(F[t]); ind = t; }
    if (mx == 0.0) break;
    const int ii = (int)(ind % (size_t)m), jj = (int)(ind / (size_t)m);
    const double gamma = 1.0 / F[ind];
    for (int i = 0; i < m; i++) An[i + (size_t)(rank - 1) * m] = gamma * F[i + (size_t)jj * m];
    for (int j = 0; j < n; j++) Bn[j + (size_t)(rank - 1) * n] = F[ii + (size_t)j * m];
    double nR = 0.0;
    for (int j = 0; j < n; j++) {
      const double t = -Bn[j + (size_t)(rank - 1) * n];
      for (int i = 0; i < m; i++) {
        F[i + (size_t)j * m] = F[i + (size_t)j * m] + t * An[i + (size_t)(rank - 1) * m];
      }
    }
    for (double v : F) nR = nR + v * v;
    nR = std::sqrt(nR);
    if (nR <= eps * nF) break;
  }
  if (rank > mn) rank = mn;
  An.resize((size_t)m * rank);
  Bn.resize((size_t)n * rank);
  A.swap(An);
  B.swap(Bn);
  return rank;
}

}  // namespace

int recompress_block(int m, int n, int k, std::vector<double> &A, std::vector<double> &B, double eps) {
  return recompress(m, n, k, A, B, eps);
}

void recompress_all(const double *hA, const int64_t *doa, const double *hB, const int64_t *dob, const int *m,
                    const int *n, int nl, double eps, std::vector<int> &rank, std::vector<std::vector<double>> &Af,
                    std::vector<std::vector<double>> &Bf) {
  Af.assign(nl, {});
  Bf.assign(nl, {});
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < nl; i++) {
    std::vector<double> a(hA + doa[i], hA + doa[i] + (int64_t)m[i] * rank[i]);
    std::vector<double> b(hB + dob[i], hB + dob[i] + (int64_t)n[i] * rank[i]);
    rank[i] = recompress(m[i], n[i], rank[i], a, b, eps);
    Af[i].swap(a);
    Bf[i].swap(b);
  }
}

int build_partition(const HostProblem &hp, int op, HostPartition &P) {
  const int n = hp.N;
  if (op == DFN_OP_N && hp.mode != DFN_MODE_II) return DFN_ERR_ARG;
  TreeBuilder tb;
  tb.nleaf = hp.nleaf;
  tb.cx.resize(n); tb.cy.resize(n); tb.perm.resize(n);
  for (int k = 0; k < n; k++) { tb.cx[k] = hp.elem[2 * k]; tb.cy[k] = hp.elem[2 * k + 1]; tb.perm[k] = k; }
  tb.make(0, n - 1);
  if (tb.fail) return DFN_ERR_ARG;
  BlockTreeBuilder bt(tb.cl, hp.eta);
  bt.make(0, 0);

  P = HostPartition();
  P.n = n;
  P.mode3 = hp.mode == DFN_MODE_III;
  P.normal = op == DFN_OP_N;
  // geometry in H order (forward_reorder, src/hmatrix/initialisation_hmat.f90:65-72)
  HGeom &g = P.g;
  g.ex = tb.cx; g.ey = tb.cy;
  for (auto *v : {&g.nlx, &g.nly, &g.nrx, &g.nry, &g.tx, &g.ty, &g.nx, &g.ny}) v->resize(n);
  for (int k = 0; k < n; k++) {
    const int s = tb.perm[k];
    g.nlx[k] = hp.node[2 * s]; g.nly[k] = hp.node[2 * s + 1];
    g.nrx[k] = hp.node[2 * (s + 1)]; g.nry[k] = hp.node[2 * (s + 1) + 1];
    g.tx[k] = hp.tang[2 * s]; g.ty[k] = hp.tang[2 * s + 1];
    g.nx[k] = hp.nrm[2 * s]; g.ny[k] = hp.nrm[2 * s + 1];
  }
  P.perm = tb.perm;
  P.nbclus = (int)tb.cl.size();
  P.nbblclus = (int)bt.blocks.size();
  for (const Cluster &c : tb.cl)
    if (c.son1 < 0) { P.leaf_r0.push_back(c.first); P.leaf_m.push_back(c.size()); }
  // leaves in preorder: dense (compute_full_matrices, H_matrix_construction.f90:555-704)
  // and admissible ones
  for (const Block &b : bt.blocks) {
    if (b.type == 0) {
      P.d_r0.push_back(tb.cl[b.row].first); P.d_m.push_back(tb.cl[b.row].size());
      P.d_c0.push_back(tb.cl[b.col].first); P.d_n.push_back(tb.cl[b.col].size());
    } else if (b.type == 1) {
      P.l_r0.push_back(tb.cl[b.row].first); P.l_m.push_back(tb.cl[b.row].size());
      P.l_c0.push_back(tb.cl[b.col].first); P.l_n.push_back(tb.cl[b.col].size());
    }
  }
  return DFN_OK;
}

void init_operator(const HostPartition &P, HostOperator &out) {
  out = HostOperator();
  out.n = P.n;
  out.perm = P.perm;
  out.nbclus = P.nbclus;
  out.nbblclus = P.nbblclus;
  out.leaf_r0 = P.leaf_r0;
  out.leaf_m = P.leaf_m;
  const int nd = (int)P.d_r0.size();
  out.d_r0 = P.d_r0; out.d_c0 = P.d_c0; out.d_m = P.d_m; out.d_n = P.d_n;
  out.d_off.resize(nd);
  int64_t off = 0;
  for (int i = 0; i < nd; i++) {
    out.d_off[i] = off;
    off += (int64_t)P.d_m[i] * P.d_n[i];
  }
  out.d_data.resize(off);
  out.l_r0 = P.l_r0; out.l_c0 = P.l_c0; out.l_m = P.l_m; out.l_n = P.l_n;
}

void pack_lowrank(std::vector<std::vector<double>> &Af, std::vector<std::vector<double>> &Bf,
                  const std::vector<int> &rank, HostOperator &out) {
  const int nl = (int)out.l_m.size();
  out.l_k = rank;
  out.l_uoff.resize(nl); out.l_voff.resize(nl);
  int64_t ou = 0, ov = 0;
  for (int i = 0; i < nl; i++) {
    out.l_uoff[i] = ou; out.l_voff[i] = ov;
    ou += (int64_t)out.l_m[i] * rank[i];
    ov += (int64_t)out.l_n[i] * rank[i];
  }
  out.l_u.resize(ou);
  out.l_v.resize(ov);
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < nl; i++) {
    std::copy(Af[i].begin(), Af[i].end(), out.l_u.begin() + out.l_uoff[i]);
    std::copy(Bf[i].begin(), Bf[i].end(), out.l_v.begin() + out.l_voff[i]);
  }
}

int build_operator(const HostProblem &hp, int op, HostOperator &out) {
  HostPartition P;
  DFN_TRY_HOST(build_partition(hp, op, P));
  init_operator(P, out);
  const HGeom &g = P.g;
  const bool mode3 = P.mode3, normal = P.normal;
  auto assemble = [&](int r0, int m, int c0, int nn, double *dst) {
    for (int j = 0; j < nn; j++)
      for (int i = 0; i < m; i++)
        dst[i + (size_t)j * m] = mode3 ? entry_III(g, r0 + i, c0 + j) : entry_II(g, r0 + i, c0 + j, normal);
  };
  const int nd = (int)out.d_r0.size();
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < nd; i++) assemble(out.d_r0[i], out.d_m[i], out.d_c0[i], out.d_n[i], &out.d_data[out.d_off[i]]);

  // admissible leaves: full-pivot ACA then recompression, largest blocks first
  const int nl = (int)out.l_m.size();
  std::vector<std::vector<double>> Af(nl), Bf(nl);
  std::vector<int> rank(nl);
  std::vector<int> order(nl);
  std::iota(order.begin(), order.end(), 0);
  auto area = [&](int i) { return (int64_t)out.l_m[i] * out.l_n[i]; };
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return area(x) > area(y); });
  int status = DFN_OK;
#pragma omp parallel for schedule(dynamic, 1)
  for (int t = 0; t < nl; t++) {
    const int i = order[t];
    const int r0 = out.l_r0[i], m = out.l_m[i], c0 = out.l_c0[i], nn = out.l_n[i];
    std::vector<double> R;
    try {
      R.resize((size_t)m * nn);
    } catch (...) {
#pragma omp atomic write
      status = DFN_ERR_NOMEM;
      continue;
    }
    assemble(r0, m, c0, nn, R.data());
    std::vector<double> A((size_t)m * hp.max_rank, 0.0), B((size_t)nn * hp.max_rank, 0.0);
    const int k = full_pivot_aca(m, nn, R, hp.eps_aca, hp.max_rank, A.data(), B.data());
    A.resize((size_t)m * k);
    B.resize((size_t)nn * k);
    rank[i] = recompress(m, nn, k, A, B, hp.eps_aca);
    Af[i].swap(A);
    Bf[i].swap(B);
  }
  if (status != DFN_OK) return status;
  pack_lowrank(Af, Bf, rank, out);
  return DFN_OK;
}

}  // namespace dfn
