// Device construction of the H-matrix operators.
//
// The partition (cluster tree, block cluster tree) is built on the host by
// build_partition; this file assembles the dense leaves
// (compute_full_matrices, src/hmatrix/H_matrix_construction.f90:555-704) and runs
// the full-pivot ACA of every admissible leaf (compute_full_ACA,
// src/hmatrix/ACA_algorithm.f90:665-826) on the GPU; the QR + SVD recompression
// (src/hmatrix/H_matrix_recompression.f90:145-301) stays on the host, on the
// factors the GPU returns.
//
// The ACA never stores a residual: R_r(i,j) = M(i,j) - a_0(i) b_0(j) - ... -
// a_{r-1}(i) b_{r-1}(j) is re-formed from the matrix entry and the factors so far,
// subtracting in the same order as the host's in-place updates, with no fused
// multiply-add, so every residual entry, pivot and factor is the same double as
// the host construction's.  Each pass over the admissible leaves (one wave per
// 64-row x 256-column tile) yields per-tile partial sums of squares and the
// first maximum of |R| in column-major order; one workgroup per leaf then sums
// the partials in a fixed order, applies the stopping rules and forms the next
// factor pair.  Only the sums of squares (used against eps) are summed in a
// different order than the host's serial loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "dfn_device.hpp"
#include "dfn_host.hpp"

#pragma clang fp contract(off)

namespace dfn {

namespace {

constexpr int kAcaRows = 64;    // rows per tile (one wave, one row per lane)
constexpr int kAcaCols = 256;   // columns per tile, in chunks of 64
constexpr int kAcaChunk = 64;

// phase timings on stderr when DFN_VERBOSE is set
struct PhaseClock {
  bool on = std::getenv("DFN_VERBOSE") != nullptr;
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  void lap(const char *what) {
    const auto now = std::chrono::steady_clock::now();
    if (on) std::fprintf(stderr, "[dfn build] %-28s %8.3f s\n", what, std::chrono::duration<double>(now - t).count());
    t = now;
  }
};

struct GeomDev {
  const double *ex, *ey, *nlx, *nly, *nrx, *nry, *tx, *ty, *nx, *ny;
};

struct DenseJob {
  int r0, c0, m, n;
  int64_t off;
};

struct AcaJob {
  int r0, c0, m, n;
  int64_t aoff, boff;  // A: m x max_rank, B: n x max_rank, column-major
};

struct AcaTile {
  int job, i0, j0, pad;
};

struct AcaPart {
  double s2, best, val;
  int64_t p;
};

struct AcaStat {
  double nA;
  int rank, done;
};

// eval_kernel_ModeIII (src/hmatrix/eval_kernel_hmatrix.f90:7-52): receiver centre
// (e1, e2) and tangent (t1, t2), source nodes (lx, ly) and (rx, ry)
__device__ __forceinline__ double entry3(double e1, double e2, double t1, double t2, double lx, double ly, double rx,
                                         double ry) {
  const double a1 = lx - e1, a2 = ly - e2, b1 = rx - e1, b2 = ry - e2;
  const double r1 = sqrt(a1 * a1 + a2 * a2), r2 = sqrt(b1 * b1 + b2 * b2);
  const double I1 = t1 / (r1 * r1) * (e1 - lx) + t2 / (r1 * r1) * (e2 - ly);
  const double I2 = t1 / (r2 * r2) * (e1 - rx) + t2 / (r2 * r2) * (e2 - ry);
  return -I1 + I2;
}

// eval_kernel_ModeII_T / _N (:57-199): receiver centre and normal (n1, n2), source
// nodes and normal (ml1, ml2)
template <bool NORMAL>
__device__ __forceinline__ double entry2(double e1, double e2, double n1, double n2, double lx, double ly, double rx,
                                         double ry, double ml1, double ml2) {
  const double exl = e1 - lx, eyl = e2 - ly, exr = e1 - rx, eyr = e2 - ry;
  const double rl = sqrt(exl * exl + eyl * eyl), rr = sqrt(exr * exr + eyr * eyr);
  const double g1l = exl / rl, g2l = eyl / rl, g1r = exr / rr, g2r = eyr / rr;
  if (!NORMAL) {
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

// Receiver data of H-order row k: centre and tangent (Mode III) or normal (Mode II).
struct Recv {
  double e1, e2, q1, q2;
};

template <int MODE>
__device__ __forceinline__ Recv recv_of(const GeomDev &g, int k) {
  if (MODE == 0) return Recv{g.ex[k], g.ey[k], g.tx[k], g.ty[k]};
  return Recv{g.ex[k], g.ey[k], g.nx[k], g.ny[k]};
}

// MODE 0: Mode III; 1: Mode II shear (T); 2: Mode II normal (N)
template <int MODE>
__device__ __forceinline__ double entry(const Recv &rv, double lx, double ly, double rx, double ry, double ml1,
                                        double ml2) {
  if (MODE == 0) return entry3(rv.e1, rv.e2, rv.q1, rv.q2, lx, ly, rx, ry);
  if (MODE == 1) return entry2<false>(rv.e1, rv.e2, rv.q1, rv.q2, lx, ly, rx, ry, ml1, ml2);
  return entry2<true>(rv.e1, rv.e2, rv.q1, rv.q2, lx, ly, rx, ry, ml1, ml2);
}

template <int MODE>
__device__ __forceinline__ double entry_at(const GeomDev &g, const Recv &rv, int l) {
  return entry<MODE>(rv, g.nlx[l], g.nly[l], g.nrx[l], g.nry[l], g.nx[l], g.ny[l]);
}

// ---------------------------------------------------------------------------
// dense leaves: one workgroup per leaf, column-major into the host layout
template <int MODE>
__global__ void __launch_bounds__(256) dense_fill_kernel(const DenseJob *__restrict__ jobs, GeomDev g,
                                                         double *__restrict__ out) {
  const DenseJob J = jobs[blockIdx.x];
  const int64_t mn = (int64_t)J.m * J.n;
  for (int64_t t = threadIdx.x; t < mn; t += blockDim.x) {
    const int i = (int)(t % J.m), j = (int)(t / J.m);
    const Recv rv = recv_of<MODE>(g, J.r0 + i);
    out[J.off + t] = entry_at<MODE>(g, rv, J.c0 + j);
  }
}

// "better" for the first maximum of |R| in column-major order: larger |v|, or the
// same |v| at a smaller linear index
__device__ __forceinline__ bool better(double b1, int64_t p1, double b0, int64_t p0) {
  return b1 > b0 || (b1 == b0 && p1 < p0);
}

// One pass: R_r over the tile (r = rk[job] factors so far), per-tile partials.
template <int MODE>
__global__ void __launch_bounds__(kAcaRows) aca_pass_kernel(const AcaTile *__restrict__ tiles,
                                                            const AcaJob *__restrict__ jobs,
                                                            const int *__restrict__ rk, const double *__restrict__ A,
                                                            const double *__restrict__ B, GeomDev g,
                                                            AcaPart *__restrict__ parts) {
  extern __shared__ double sb[];  // [kMaxRank][kAcaChunk]: b_f(j) of the chunk
  __shared__ double slx[kAcaChunk], sly[kAcaChunk], srx[kAcaChunk], sry[kAcaChunk], smx[kAcaChunk],
      smy[kAcaChunk];
  const AcaTile t = tiles[blockIdx.x];
  const AcaJob J = jobs[t.job];
  const int r = rk[t.job];
  const int lane = threadIdx.x;
  const int i = t.i0 + lane;
  const bool row_ok = i < J.m;
  Recv rv{0.0, 0.0, 0.0, 0.0};
  if (row_ok) rv = recv_of<MODE>(g, J.r0 + i);
  double av[kMaxRank];
#pragma unroll
  for (int f = 0; f < kMaxRank; f++) av[f] = (f < r && row_ok) ? A[J.aoff + (int64_t)f * J.m + i] : 0.0;
  double s2 = 0.0, best = -1.0, val = 0.0;
  int64_t p = 0;
  const int jend = min(t.j0 + kAcaCols, J.n);
  for (int c0 = t.j0; c0 < jend; c0 += kAcaChunk) {
    const int cols = min(kAcaChunk, jend - c0);
    __syncthreads();
    if (lane < cols) {
      const int l = J.c0 + c0 + lane;
      slx[lane] = g.nlx[l]; sly[lane] = g.nly[l]; srx[lane] = g.nrx[l]; sry[lane] = g.nry[l];
      smx[lane] = g.nx[l]; smy[lane] = g.ny[l];
      for (int f = 0; f < r; f++) sb[f * kAcaChunk + lane] = B[J.boff + (int64_t)f * J.n + c0 + lane];
    }
    __syncthreads();
    if (row_ok) {
      for (int q = 0; q < cols; q++) {
        double v = entry<MODE>(rv, slx[q], sly[q], srx[q], sry[q], smx[q], smy[q]);
#pragma unroll
        for (int f = 0; f < kMaxRank; f++)
          if (f < r) v = v - av[f] * sb[f * kAcaChunk + q];
        s2 = s2 + v * v;
        const double a = fabs(v);
        if (a > best) { best = a; val = v; p = (int64_t)(c0 + q) * J.m + i; }
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s2 += __shfl_xor(s2, off, 64);
    const double ob = __shfl_xor(best, off, 64), ov = __shfl_xor(val, off, 64);
    const int64_t op = __shfl_xor(p, off, 64);
    if (better(ob, op, best, p)) { best = ob; val = ov; p = op; }
  }
  if (lane == 0) parts[blockIdx.x] = AcaPart{s2, best, val, p};
}

// One workgroup per active leaf: fixed-order reduction of the pass partials, the
// stopping rules of compute_full_ACA, then the next factor pair
// a_r = R_r(:, jp), b_r = R_r(ip, :) / R_r(ip, jp).
template <int MODE>
__global__ void __launch_bounds__(256) aca_pivot_kernel(const int *__restrict__ active,
                                                        const int64_t *__restrict__ tstart,
                                                        const int *__restrict__ tcount, const AcaJob *__restrict__ jobs,
                                                        AcaStat *__restrict__ st, int *__restrict__ rk,
                                                        const AcaPart *__restrict__ parts, double *__restrict__ A,
                                                        double *__restrict__ B, GeomDev g, double eps, int max_rank,
                                                        double piv_tol) {
  __shared__ double rs2[256], rbest[256], rval[256];
  __shared__ int64_t rp[256];
  __shared__ double fa[kMaxRank], fb[kMaxRank];
  __shared__ int go;
  __shared__ double piv;
  __shared__ int ip, jp;
  const int b = active[blockIdx.x];
  const AcaJob J = jobs[b];
  const int r = rk[b];
  const int tid = threadIdx.x;
  double s2 = 0.0, best = -1.0, val = 0.0;
  int64_t p = 0;
  const int64_t t0 = tstart[blockIdx.x];
  const int tn = tcount[blockIdx.x];
  for (int t = tid; t < tn; t += 256) {
    const AcaPart q = parts[t0 + t];
    s2 = s2 + q.s2;
    if (better(q.best, q.p, best, p)) { best = q.best; val = q.val; p = q.p; }
  }
  rs2[tid] = s2; rbest[tid] = best; rval[tid] = val; rp[tid] = p;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) {
      rs2[tid] = rs2[tid] + rs2[tid + h];
      if (better(rbest[tid + h], rp[tid + h], rbest[tid], rp[tid])) {
        rbest[tid] = rbest[tid + h]; rval[tid] = rval[tid + h]; rp[tid] = rp[tid + h];
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    AcaStat S = st[b];
    const int rmax = min(J.m, J.n);
    bool done = false;
    int rank = 0;
    if (r == 0) {
      S.nA = sqrt(rs2[0]);
    } else {
      const double nR = sqrt(rs2[0]);
      if (nR <= eps * S.nA) { done = true; rank = r; }
      else if (r + 1 > max_rank) { done = true; rank = r; }
      else if (r + 1 > rmax) { done = true; rank = rmax; }  // the loop over r ran out
    }
    if (!done && fabs(rval[0]) <= piv_tol) {  // pivot too small at iteration r+1
      done = true;
      rank = r > 0 ? r : min(1, rmax);
    }
    go = 0;
    if (done) {
      S.done = 1;
      S.rank = rank;
    } else {
      go = 1;
      piv = rval[0];
      ip = (int)(rp[0] % J.m);
      jp = (int)(rp[0] / J.m);
    }
    st[b] = S;
  }
  __syncthreads();
  if (!go) return;
  if (tid < r) {
    fa[tid] = A[J.aoff + (int64_t)tid * J.m + ip];  // a_f(ip)
    fb[tid] = B[J.boff + (int64_t)tid * J.n + jp];  // b_f(jp)
  }
  __syncthreads();
  for (int i = tid; i < J.m; i += 256) {  // a_r = R_r(:, jp)
    const Recv rv = recv_of<MODE>(g, J.r0 + i);
    double v = entry_at<MODE>(g, rv, J.c0 + jp);
    for (int f = 0; f < r; f++) v = v - A[J.aoff + (int64_t)f * J.m + i] * fb[f];
    A[J.aoff + (int64_t)r * J.m + i] = v;
  }
  const Recv rp_ = recv_of<MODE>(g, J.r0 + ip);
  for (int j = tid; j < J.n; j += 256) {  // b_r = R_r(ip, :) / pivot
    double v = entry_at<MODE>(g, rp_, J.c0 + j);
    for (int f = 0; f < r; f++) v = v - fa[f] * B[J.boff + (int64_t)f * J.n + j];
    B[J.boff + (int64_t)r * J.n + j] = v / piv;
  }
  if (tid == 0) rk[b] = r + 1;
}

// copies the first k columns of every leaf's A (or B) into a packed buffer
__global__ void __launch_bounds__(256) aca_pack_kernel(const double *__restrict__ src, const int64_t *__restrict__ soff,
                                                       const int64_t *__restrict__ doff,
                                                       const int64_t *__restrict__ cnt, double *__restrict__ dst) {
  const int b = blockIdx.x;
  const int64_t n = cnt[b], s = soff[b], d = doff[b];
  for (int64_t t = threadIdx.x; t < n; t += blockDim.x) dst[d + t] = src[s + t];
}

// ---------------------------------------------------------------------------
struct DBuf {
  void *p = nullptr;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t bytes) {
    if (p) (void)hipFree(p);
    p = nullptr;
    if (bytes == 0) bytes = 8;
    return hipMalloc(&p, bytes) == hipSuccess ? DFN_OK : DFN_ERR_NOMEM;
  }
  template <class T>
  int upload(const std::vector<T> &v, hipStream_t s) {
    DFN_TRY_HOST(alloc(sizeof(T) * v.size()));
    if (v.empty()) return DFN_OK;
    return hipMemcpyAsync(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s) == hipSuccess ? DFN_OK
                                                                                                      : DFN_ERR_HIP;
  }
  template <class T>
  T *as() const {
    return (T *)p;
  }
};

#define HB_HIP(x)                                \
  do {                                           \
    if ((x) != hipSuccess) return DFN_ERR_HIP; \
  } while (0)

template <int MODE>
int run_device_build(const HostProblem &hp, const HostPartition &P, HostOperator &out, hipStream_t s,
                     PhaseClock &clk) {
  const int n = P.n;
  DBuf geo[10];
  const std::vector<double> *gv[10] = {&P.g.ex, &P.g.ey, &P.g.nlx, &P.g.nly, &P.g.nrx,
                                       &P.g.nry, &P.g.tx, &P.g.ty,  &P.g.nx,  &P.g.ny};
  for (int q = 0; q < 10; q++) DFN_TRY_HOST(geo[q].upload(*gv[q], s));
  GeomDev g{geo[0].as<double>(), geo[1].as<double>(), geo[2].as<double>(), geo[3].as<double>(),
            geo[4].as<double>(), geo[5].as<double>(), geo[6].as<double>(), geo[7].as<double>(),
            geo[8].as<double>(), geo[9].as<double>()};
  (void)n;

  // dense leaves
  const int nd = (int)out.d_m.size();
  if (nd > 0) {
    std::vector<DenseJob> dj(nd);
    for (int i = 0; i < nd; i++) dj[i] = DenseJob{out.d_r0[i], out.d_c0[i], out.d_m[i], out.d_n[i], out.d_off[i]};
    DBuf djb, dd;
    DFN_TRY_HOST(djb.upload(dj, s));
    DFN_TRY_HOST(dd.alloc(sizeof(double) * out.d_data.size()));
    hipLaunchKernelGGL(dense_fill_kernel<MODE>, dim3(nd), dim3(256), 0, s, djb.as<DenseJob>(), g, dd.as<double>());
    HB_HIP(hipGetLastError());
    HB_HIP(hipMemcpyAsync(out.d_data.data(), dd.p, sizeof(double) * out.d_data.size(), hipMemcpyDeviceToHost, s));
    HB_HIP(hipStreamSynchronize(s));
  }
  clk.lap("dense leaves (GPU + copy)");

  // admissible leaves
  const int nl = (int)out.l_m.size();
  std::vector<int> rank(nl, 0);
  std::vector<std::vector<double>> Af(nl), Bf(nl);
  if (nl > 0) {
    const int mr = hp.max_rank;
    std::vector<AcaJob> jobs(nl);
    int64_t ao = 0, bo = 0;
    for (int i = 0; i < nl; i++) {
      jobs[i] = AcaJob{out.l_r0[i], out.l_c0[i], out.l_m[i], out.l_n[i], ao, bo};
      ao += (int64_t)out.l_m[i] * mr;
      bo += (int64_t)out.l_n[i] * mr;
    }
    DBuf jb, Ab, Bb, stb, rkb;
    DFN_TRY_HOST(jb.upload(jobs, s));
    DFN_TRY_HOST(Ab.alloc(sizeof(double) * ao));
    DFN_TRY_HOST(Bb.alloc(sizeof(double) * bo));
    HB_HIP(hipMemsetAsync(Ab.p, 0, sizeof(double) * ao, s));
    HB_HIP(hipMemsetAsync(Bb.p, 0, sizeof(double) * bo, s));
    std::vector<AcaStat> st0(nl, AcaStat{0.0, 0, 0});
    DFN_TRY_HOST(stb.upload(st0, s));
    std::vector<int> rk0(nl, 0);
    DFN_TRY_HOST(rkb.upload(rk0, s));
    std::vector<int> active(nl);
    std::iota(active.begin(), active.end(), 0);
    DBuf tb, pb, ab, tsb, tcb;
    const size_t lds = sizeof(double) * kMaxRank * kAcaChunk;
    const double piv_tol = (double)1.0e-14f;  // ACA_algorithm.f90: pivot tolerance, a single-precision literal
    std::vector<AcaStat> stv(nl);
    for (int pass = 0; !active.empty(); pass++) {
      // tiles of the active leaves, grouped by leaf
      std::vector<AcaTile> tiles;
      std::vector<int64_t> tstart(active.size());
      std::vector<int> tcount(active.size());
      for (size_t a = 0; a < active.size(); a++) {
        const AcaJob &J = jobs[active[a]];
        tstart[a] = (int64_t)tiles.size();
        for (int j0 = 0; j0 < J.n; j0 += kAcaCols)
          for (int i0 = 0; i0 < J.m; i0 += kAcaRows) tiles.push_back(AcaTile{active[a], i0, j0, 0});
        tcount[a] = (int)((int64_t)tiles.size() - tstart[a]);
      }
      DFN_TRY_HOST(tb.upload(tiles, s));
      DFN_TRY_HOST(pb.alloc(sizeof(AcaPart) * tiles.size()));
      DFN_TRY_HOST(ab.upload(active, s));
      DFN_TRY_HOST(tsb.upload(tstart, s));
      DFN_TRY_HOST(tcb.upload(tcount, s));
      hipLaunchKernelGGL(aca_pass_kernel<MODE>, dim3((unsigned)tiles.size()), dim3(kAcaRows), lds, s,
                         tb.as<AcaTile>(), jb.as<AcaJob>(), rkb.as<int>(), Ab.as<double>(), Bb.as<double>(), g,
                         pb.as<AcaPart>());
      HB_HIP(hipGetLastError());
      hipLaunchKernelGGL(aca_pivot_kernel<MODE>, dim3((unsigned)active.size()), dim3(256), 0, s, ab.as<int>(),
                         tsb.as<int64_t>(), tcb.as<int>(), jb.as<AcaJob>(), stb.as<AcaStat>(), rkb.as<int>(),
                         pb.as<AcaPart>(), Ab.as<double>(), Bb.as<double>(), g, hp.eps_aca, mr, piv_tol);
      HB_HIP(hipGetLastError());
      HB_HIP(hipMemcpyAsync(stv.data(), stb.p, sizeof(AcaStat) * nl, hipMemcpyDeviceToHost, s));
      HB_HIP(hipStreamSynchronize(s));
      std::vector<int> next;
      next.reserve(active.size());
      for (int b : active)
        if (!stv[b].done) next.push_back(b);
      active.swap(next);
      if (pass > mr + 1) return DFN_ERR_STATE;  // every leaf stops by max_rank passes
    }
    clk.lap("ACA passes (GPU)");
    // the first rank columns of every leaf's factors, packed, back to the host
    std::vector<int64_t> soa(nl), sob(nl), doa(nl), dob(nl), ca(nl), cb(nl);
    int64_t pa = 0, pb2 = 0;
    for (int i = 0; i < nl; i++) {
      rank[i] = stv[i].rank;
      soa[i] = jobs[i].aoff; sob[i] = jobs[i].boff;
      ca[i] = (int64_t)jobs[i].m * rank[i]; cb[i] = (int64_t)jobs[i].n * rank[i];
      doa[i] = pa; dob[i] = pb2;
      pa += ca[i]; pb2 += cb[i];
    }
    DBuf pkA, pkB, b1, b2, b3, b4, b5, b6;
    DFN_TRY_HOST(pkA.alloc(sizeof(double) * pa));
    DFN_TRY_HOST(pkB.alloc(sizeof(double) * pb2));
    DFN_TRY_HOST(b1.upload(soa, s)); DFN_TRY_HOST(b2.upload(doa, s)); DFN_TRY_HOST(b3.upload(ca, s));
    DFN_TRY_HOST(b4.upload(sob, s)); DFN_TRY_HOST(b5.upload(dob, s)); DFN_TRY_HOST(b6.upload(cb, s));
    hipLaunchKernelGGL(aca_pack_kernel, dim3(nl), dim3(256), 0, s, Ab.as<double>(), b1.as<int64_t>(),
                       b2.as<int64_t>(), b3.as<int64_t>(), pkA.as<double>());
    hipLaunchKernelGGL(aca_pack_kernel, dim3(nl), dim3(256), 0, s, Bb.as<double>(), b4.as<int64_t>(),
                       b5.as<int64_t>(), b6.as<int64_t>(), pkB.as<double>());
    HB_HIP(hipGetLastError());
    std::vector<double> hA(pa), hB(pb2);
    HB_HIP(hipMemcpyAsync(hA.data(), pkA.p, sizeof(double) * pa, hipMemcpyDeviceToHost, s));
    HB_HIP(hipMemcpyAsync(hB.data(), pkB.p, sizeof(double) * pb2, hipMemcpyDeviceToHost, s));
    HB_HIP(hipStreamSynchronize(s));
    clk.lap("factors back to host");
    // recompression on the host
    recompress_all(hA.data(), doa.data(), hB.data(), dob.data(), out.l_m.data(), out.l_n.data(), nl, hp.eps_aca,
                   rank, Af, Bf);
    clk.lap("recompression (host)");
  }
  pack_lowrank(Af, Bf, rank, out);
  clk.lap("pack");
  return DFN_OK;
}

}  // namespace

int build_operator_device(const HostProblem &hp, int op, HostOperator &out, void *stream) {
  if (hp.max_rank > kMaxRank || hp.max_rank < 1) return build_operator(hp, op, out);
  PhaseClock clk;
  HostPartition P;
  DFN_TRY_HOST(build_partition(hp, op, P));
  init_operator(P, out);
  clk.lap("partition (host)");
  hipStream_t s = (hipStream_t)stream;
  if (P.mode3) return run_device_build<0>(hp, P, out, s, clk);
  if (!P.normal) return run_device_build<1>(hp, P, out, s, clk);
  return run_device_build<2>(hp, P, out, s, clk);
}

}  // namespace dfn
