// Shared by the single-network engine and the ensemble engine: error macros,
// owning device buffers and the device layout of the stress-transfer operators
// (two-phase H-matvec task lists, see kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "dfn_device.hpp"
#include "dfn_host.hpp"

namespace dfn {

#define DFN_HIP(expr)                                       \
  do {                                                      \
    hipError_t e_ = (expr);                                 \
    if (e_ != hipSuccess) {                                 \
      std::fprintf(stderr, "dfn: %s failed: %s\n", #expr,   \
                   hipGetErrorString(e_));                  \
      return DFN_ERR_HIP;                                   \
    }                                                       \
  } while (0)

#define DFN_TRY(expr)           \
  do {                          \
    int s_ = (expr);            \
    if (s_ != DFN_OK) return s_; \
  } while (0)

// owning device allocation
struct DMem {
  void *p = nullptr;
  size_t bytes = 0;
  DMem() = default;
  DMem(const DMem &) = delete;
  DMem &operator=(const DMem &) = delete;
  ~DMem() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  int alloc(size_t nbytes) {
    release();
    if (nbytes == 0) nbytes = 8;
    if (hipMalloc(&p, nbytes) != hipSuccess) { p = nullptr; return DFN_ERR_NOMEM; }
    bytes = nbytes;
    return DFN_OK;
  }
  template <class T>
  int upload(const T *v, size_t count) {
    DFN_TRY(alloc(count * sizeof(T)));
    if (count) DFN_HIP(hipMemcpy(p, v, count * sizeof(T), hipMemcpyHostToDevice));
    return DFN_OK;
  }
  template <class T>
  int upload(const std::vector<T> &v) {
    return upload(v.data(), v.size());
  }
  template <class T>
  T *as() const { return static_cast<T *>(p); }
};

// Device layout of the stress-transfer operator(s) of one problem (1 in Mode III,
// T and N in Mode II, sharing the permutation and the input vector).
struct DevOps {
  int n = 0, nops = 0;
  bool fp32 = false;  // phase 1 reads FP32 V factors (and the FP32 copy of x_h)
  DMem perm, vtasks, vdata, zred, zpart, ptasks, pdata, widx, cscale, w, ypart, fin_first, fin_idx;
  // phase-2 panel sets: set A (ptasks, pdata, widx, cscale) in FP32 when pa_fp32, else
  // FP64; with FP64 near-field storage beside FP32 low-rank factors (DFN_PREC_MIXED),
  // set A holds the near-field (dense) columns in FP64 and set B the U columns in FP32
  bool pa_fp32 = false;
  DMem ptasks_b, pdata_b, widx_b, cscale_b;
  // split: set A holds the near-field entries as FP32 values (pdata) and 8-bit
  // residuals (pdata_lo, the same entry offsets), cscale the residual's scale of
  // each task (repeated over the task's columns)
  bool split = false;
  DMem pdata_lo;
  int npt_b = 0;
  DMem xf;  // FP32 operators: FP32 copy of x_h read by phase 1 (n + 4 floats, the tail zero)
  // FP32 storage read into FP64 products and sums (DFN_PREC_FP32_STORE, and the
  // low-rank factors of DFN_PREC_MIXED)
  bool acc64 = false;
  // DFN_PREC_MIXED: phase 1 and the FP64 near-field panels (set A, which read only
  // x_h) may run as one launch (launch_vtx_near); near_ok says the layout allows it,
  // fused that the autotuner chose it
  bool near_ok = false, fused = false;
  int vmode() const { return fp32 ? (acc64 ? kStoreF32Acc64 : kStoreF32) : vnt ? kStoreF64NT : kStoreF64; }
  int pmode_a() const { return pa_fp32 ? (acc64 ? kStoreF32Acc64 : kStoreF32) : kStoreF64; }
  int pmode_b() const { return acc64 ? kStoreF32Acc64 : kStoreF32; }
  int nvt = 0, nzr = 0, npt = 0;
  int kmax = kMaxRank;  // largest rank of any phase-1 task
  int kc = kMaxRank;    // rank ceiling of the phase-1 kernel instance in use (16, 20 or 32, >= kmax)
  double alg_bytes = 0;    // algorithmic HBM bytes per application (all operators, all phases)
  double panel_bytes = 0;  // algorithmic bytes of the phase-2 panel kernel (dense + U entries, y partials)
  // algorithmic bytes of launch_vtx_near (FP32 V entries, FP64 near-field entries,
  // x_h, z and near-field row partials written)
  double fused_bytes = 0;
  // the dominant launch of an evaluation and its algorithmic bytes: the fused
  // launch when fused, else the phase-2 panels
  double dom_bytes() const { return fused ? fused_bytes : panel_bytes; }
  int64_t S_total = 0;   // matrix entries streamed per application
  int64_t S_panel = 0;   // of which in the phase-2 panels (dense and U entries)
  // launch shapes picked by autotune (none changes a sum): phase-1 tasks per
  // workgroup, phase-1 task order (0 as built, 1 grouped by rank, 2 largest tasks
  // first), panel rows per wave step, panel task order (0 as built, 1 largest
  // first) and rows per panel task (0: whole panels); and the matvec time (ms) of
  // the default and of the chosen shapes
  int vtx_waves = 4, vtx_order = 0, panel_rows = 2, panel_order = 0, panel_split = 0;
  double tune_ms_default = 0, tune_ms = 0;
  // phase-1 wave items (vpair): short tasks of one rank and network two to a wave
  bool vpair = false;
  DMem vitems;
  int nvi = 0;
  // FP64 phase-1 factors read with non-temporal loads (the same values)
  bool vnt = false;
};

// Phase-1 wave items over the task list in its launch order: each task of at most 32
// row vectors (kvec rows each) waits for the next such task of its rank in the same
// network and shares a wave with it; the other tasks keep a wave each.
inline std::vector<VItem> vtx_items(const std::vector<VTask> &v, int kvec) {
  std::vector<VItem> items;
  items.reserve(v.size());
  int pending[kMaxRank + 1];
  std::fill_n(pending, kMaxRank + 1, -1);
  int net = -1;
  auto flush = [&]() {
    for (int k = 0; k <= kMaxRank; k++)
      if (pending[k] >= 0) {
        items.push_back({pending[k], -1});
        pending[k] = -1;
      }
  };
  for (int i = 0; i < (int)v.size(); i++) {
    const VTask &t = v[i];
    if (t.len > 32 * kvec || t.k < 1 || t.k > kMaxRank) {
      items.push_back({i, -1});
      continue;
    }
    if (t.net != net) {
      flush();
      net = t.net;
    }
    if (pending[t.k] >= 0) {
      items.push_back({pending[t.k], i});
      pending[t.k] = -1;
    } else {
      pending[t.k] = i;
    }
  }
  flush();
  return items;
}

// Phase-2 schedule: each panel task cut into tasks of at most `split` rows (0: not
// cut; a row's sum is computed the same way whichever task holds it), the largest
// tasks first, so that the last workgroups to start are short ones.
inline std::vector<PTask> panel_schedule(const std::vector<PTask> &pt, int split) {
  std::vector<PTask> out;
  for (const PTask &t : pt)
    for (int r0 = 0; r0 < t.m; r0 += (split > 0 ? split : t.m)) {
      PTask s = t;
      s.m = split > 0 ? std::min(split, t.m - r0) : t.m;
      s.poff = t.poff + (int64_t)r0 * t.L;
      s.yp = t.yp + r0;
      out.push_back(s);
    }
  std::stable_sort(out.begin(), out.end(),
                   [](const PTask &x, const PTask &y) { return (int64_t)x.m * x.L > (int64_t)y.m * y.L; });
  return out;
}

// Multi-RHS stress evaluation with an FP64 layout (d.fp32 false, one panel set)
inline void launch_mrhs(const DevOps &d, const double *X, int nrhs, int R, double *W, double *zpart, double *ypart,
                        double *Yh, double scale, double *yT, double *yN, hipStream_t s, hipEvent_t *ev = nullptr) {
  MrhsLayout m;
  m.vt = d.vtasks.as<VTask>(); m.nvt = d.nvt; m.vdata = static_cast<const double *>(d.vdata.p);
  m.zr = d.zred.as<ZRed>(); m.nzr = d.nzr;
  m.pt = d.ptasks.as<PTask>(); m.npt = d.npt; m.pdata = static_cast<const double *>(d.pdata.p);
  m.widx = d.widx.as<int>(); m.fin_first = d.fin_first.as<int>(); m.fin_idx = d.fin_idx.as<int>();
  m.perm = d.perm.as<int>(); m.n = d.n; m.nops = d.nops;
  launch_mrhs(m, X, nrhs, R, W, zpart, ypart, Yh, scale, yT, yN, s, ev);
}

// Phase 2 of a stress evaluation: every panel set of the layout (rows: panel rows
// per wave step; active as for launch_panel)
inline void launch_panels(const DevOps &d, const double *w, hipStream_t s, int rows, const int *active = nullptr) {
  if (d.split)
    launch_panel_split(d.ptasks.as<PTask>(), d.npt, d.pdata.as<float>(), d.pdata_lo.as<signed char>(), d.widx.as<int>(),
                       d.cscale.as<float>(), w, d.ypart.as<double>(), s, rows, active);
  else
    launch_panel(d.ptasks.as<PTask>(), d.npt, d.pdata.p, d.pmode_a(), d.widx.as<int>(), d.cscale.as<float>(), w,
                 d.ypart.as<double>(), s, rows, active);
  if (d.npt_b > 0)
    launch_panel(d.ptasks_b.as<PTask>(), d.npt_b, d.pdata_b.p, d.pmode_b(), d.widx_b.as<int>(), d.cscale_b.as<float>(), w,
                 d.ypart.as<double>(), s, rows, active);
}

// Phase 1, the z reduction and phase 2 of one stress evaluation (x_h already
// gathered into w); kt (may be null) gets an event pair around the dominant launch
inline void launch_phases(const DevOps &d, double *w, hipStream_t s, int waves, int rows, const int *active,
                          hipEvent_t *kt) {
  if (d.fused) {
    if (kt) (void)hipEventRecord(kt[0], s);
    launch_vtx_near(d.vtasks.as<VTask>(), d.nvt, static_cast<const float *>(d.vdata.p), d.ptasks.as<PTask>(), d.npt,
                    static_cast<const double *>(d.pdata.p), d.widx.as<int>(), w, d.zpart.as<double>(),
                    d.ypart.as<double>(), s, rows, active, d.kc);
    if (kt) (void)hipEventRecord(kt[1], s);
    launch_zreduce(d.zred.as<ZRed>(), d.nzr, d.zpart.as<double>(), w, s);
    launch_panel(d.ptasks_b.as<PTask>(), d.npt_b, d.pdata_b.p, d.pmode_b(), d.widx_b.as<int>(),
                 d.cscale_b.as<float>(), w, d.ypart.as<double>(), s, rows, active);
    return;
  }
  launch_vtx(d.vtasks.as<VTask>(), d.nvt, d.vdata.p, d.vmode(), w, d.zpart.as<double>(), s, waves, d.xf.as<float>(),
             active, d.kc, d.vpair ? d.vitems.as<VItem>() : nullptr, d.nvi);
  launch_zreduce(d.zred.as<ZRed>(), d.nzr, d.zpart.as<double>(), w, s);
  if (kt) (void)hipEventRecord(kt[0], s);
  launch_panels(d, w, s, rows, active);
  if (kt) (void)hipEventRecord(kt[1], s);
}

// Autotuner of the stress-evaluation launch shapes, run once per operator layout on
// the operator itself: phase-1 tasks (waves) per workgroup {2, 4, 8}; the phase-1
// task order (as built; grouped by rank bucket with the largest ranks first, so a
// workgroup's waves share one code path; or the largest tasks, rows x rank, first);
// short phase-1 tasks one per wave or two (same rank and network) to a wave;
// 2 or 4 panel rows per wave step; the phase-2 schedule (as built, or largest
// panels first, whole or cut into tasks of 32 or 16 rows); the phase-1 kernel
// instance (the smallest rank ceiling that covers the layout, or 32); for a
// mixed-precision layout, phase 1 and the near-field panels in one launch or in
// two.  Every candidate
// computes every sum in the same order, so the choice never changes a result.
// Each candidate is timed over phase 1, the z reduction and phase 2 (HIP events,
// 2 warm-up + 8 timed applications); a candidate replaces the incumbent only if it
// is 1 % faster (the phase-1 wave items: 0.5 % faster over phase 1 alone).  DFN_AUTOTUNE=alt takes non-default shapes (8 waves, largest
// phase-1 tasks first, two short tasks to a wave, 4 rows, 16-row panel tasks)
// without timing, so that a test
// can check that the shapes agree.
inline int autotune(DevOps &d, const std::vector<VTask> &vt, const std::vector<PTask> &pt,
                    const std::vector<PTask> &pt_b) {
  hipEvent_t e0, e1;
  DFN_HIP(hipEventCreate(&e0));
  DFN_HIP(hipEventCreate(&e1));
  double *w = d.w.as<double>();
  DFN_HIP(hipMemset(w, 0, d.w.bytes));
  auto measure = [&](int waves, int rows, double &ms) -> int {
    for (int rep = 0; rep < 10; rep++) {
      if (rep == 2) DFN_HIP(hipEventRecord(e0, nullptr));
      launch_phases(d, w, nullptr, waves, rows, nullptr, nullptr);
    }
    DFN_HIP(hipEventRecord(e1, nullptr));
    DFN_HIP(hipEventSynchronize(e1));
    float t = 0;
    DFN_HIP(hipEventElapsedTime(&t, e0, e1));
    ms = t / 8.0;
    return DFN_OK;
  };
  auto measure_p1 = [&](int waves, double &ms) -> int {  // phase 1 alone
    for (int rep = 0; rep < 10; rep++) {
      if (rep == 2) DFN_HIP(hipEventRecord(e0, nullptr));
      launch_vtx(d.vtasks.as<VTask>(), d.nvt, d.vdata.p, d.vmode(), w, d.zpart.as<double>(), nullptr, waves,
                 d.xf.as<float>(), nullptr, d.kc, d.vpair ? d.vitems.as<VItem>() : nullptr, d.nvi);
    }
    DFN_HIP(hipEventRecord(e1, nullptr));
    DFN_HIP(hipEventSynchronize(e1));
    float t = 0;
    DFN_HIP(hipEventElapsedTime(&t, e0, e1));
    ms = t / 8.0;
    return DFN_OK;
  };
  std::vector<VTask> vcur;  // the phase-1 task list as uploaded
  auto vorder = [&](int order) -> int {
    vcur = vt;
    if (order == 1)
      std::stable_sort(vcur.begin(), vcur.end(), [](const VTask &x, const VTask &y) { return x.k > y.k; });
    else if (order == 2)
      std::stable_sort(vcur.begin(), vcur.end(),
                       [](const VTask &x, const VTask &y) { return x.len * x.k > y.len * y.k; });
    d.vpair = false;
    return d.vtasks.upload(vcur);
  };
  auto vpairs = [&](bool on) -> int {  // short tasks two to a wave, over the current order
    d.vpair = false;
    if (!on) return DFN_OK;
    const std::vector<VItem> items = vtx_items(vcur, d.fp32 ? 4 : 2);
    DFN_TRY(d.vitems.upload(items));
    d.nvi = (int)items.size();
    d.vpair = true;
    return DFN_OK;
  };
  auto porder = [&](int order, int split) -> int {  // order 0: as built
    const std::vector<PTask> p = order ? panel_schedule(pt, split) : pt;
    d.npt = (int)p.size();
    DFN_TRY(d.ptasks.upload(p));
    if (pt_b.empty()) return DFN_OK;
    const std::vector<PTask> q = order ? panel_schedule(pt_b, split) : pt_b;
    d.npt_b = (int)q.size();
    return d.ptasks_b.upload(q);
  };
  int waves = 4, order = 0, rows = 2, pord = 0, split = 0;
  bool pair = false;
  const char *mode = std::getenv("DFN_AUTOTUNE");
  const bool alt = mode && std::string(mode) == "alt";
  double best = 0, t = 0;
  if (alt) {
    waves = 8; order = 2; rows = 4; pord = 1; split = 16; pair = true;
    d.vnt = !d.fp32;
    d.kc = kMaxRank;  // the rank-32 instance (the single launch is exercised by DFN_FUSE_NEAR=1)
    DFN_TRY(vorder(order));
    DFN_TRY(vpairs(pair));
    DFN_TRY(porder(pord, split));
  } else {
    DFN_TRY(measure(waves, rows, best));
    d.tune_ms_default = best;
    for (int cand : {2, 8}) {
      DFN_TRY(measure(cand, rows, t));
      if (t < 0.99 * best) { best = t; waves = cand; }
    }
    for (int cand : {1, 2}) {
      DFN_TRY(vorder(cand));
      DFN_TRY(measure(waves, rows, t));
      if (t < 0.99 * best) { best = t; order = cand; }
    }
    DFN_TRY(vorder(order));
    {  // wave items, judged on phase 1 alone (3-5 % there is under 1 % of the whole)
      double t1 = 0;
      DFN_TRY(measure_p1(waves, t));
      DFN_TRY(vpairs(true));
      DFN_TRY(measure_p1(waves, t1));
      pair = t1 < 0.995 * t;
      DFN_TRY(vpairs(pair));
      if (pair) DFN_TRY(measure(waves, rows, best));
    }
    if (!d.fp32) {  // non-temporal FP64 factor loads, judged on the whole product
      d.vnt = true;
      DFN_TRY(measure(waves, rows, t));
      if (t < 0.99 * best) best = t;
      else d.vnt = false;
    }
    DFN_TRY(measure(waves, 4, t));
    if (t < 0.99 * best) { best = t; rows = 4; }
    for (int cand : {0, 32, 16}) {
      DFN_TRY(porder(1, cand));
      DFN_TRY(measure(waves, rows, t));
      if (t < 0.99 * best) { best = t; pord = 1; split = cand; }
    }
    DFN_TRY(porder(pord, split));
    if (d.kc < kMaxRank) {  // the rank-32 instance: fewer waves resident, more loads in flight each
      const int kc = d.kc;
      d.kc = kMaxRank;
      DFN_TRY(measure(waves, rows, t));
      if (t < 0.99 * best) best = t;
      else d.kc = kc;
    }
    if (d.near_ok && d.kc <= 20) {
      d.fused = true;
      DFN_TRY(measure(waves, rows, t));
      if (t < 0.99 * best) best = t;
      else d.fused = false;
    }
  }
  // DFN_VTX_PAIRS=1 / 0 takes / refuses the phase-1 wave items whatever their timing
  const char *pe = std::getenv("DFN_VTX_PAIRS");
  if (pe && (pe[0] == '0' || pe[0] == '1')) DFN_TRY(vpairs(pe[0] == '1'));
  d.vtx_waves = waves;
  d.vtx_order = order;
  d.panel_rows = rows;
  d.panel_order = pord;
  d.panel_split = split;
  d.tune_ms = best;
  DFN_HIP(hipEventDestroy(e0));
  DFN_HIP(hipEventDestroy(e1));
  return DFN_OK;
}

// f(i) for i in [0, n) on up to 16 host threads (items handed out one at a time,
// so each item should be a block- or panel-sized piece of work; f must not throw).
// If a thread cannot be started, the threads that did start and the caller share
// the work.
template <class F>
void parallel_for(int64_t n, F f) {
  const int nt = (int)std::min<int64_t>(n, std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
  std::atomic<int64_t> next{0};
  auto work = [&] {
    for (int64_t i; (i = next.fetch_add(1)) < n;) f(i);
  };
  std::vector<std::thread> pool;
  pool.reserve(nt);
  for (int t = 1; t < nt; t++) {
    try {
      pool.emplace_back(work);
    } catch (...) {
      break;
    }
  }
  work();
  for (auto &t : pool) t.join();
}

// power-of-two scale 2^e >= max|v| (1 for an all-zero block): dividing by it is exact
inline double pow2_scale(const double *v, size_t n) {
  double m = 0.0;
  for (size_t i = 0; i < n; i++) m = std::fmax(m, std::fabs(v[i]));
  if (!(m > 0.0) || !std::isfinite(m)) return 1.0;
  int e = 0;
  (void)std::frexp(m, &e);
  return std::ldexp(1.0, e);
}

// hnet (ensembles, may be null): the network of every H position; each phase-1 and
// phase-2 task records the network its block or panel belongs to, so that the
// kernels can skip the tasks of networks that sit out an evaluation.
// T: storage of the low-rank factors (and of the near-field blocks unless dense64,
// which keeps them in their own panel set: in FP64, or with split as an FP32 value
// and an 8-bit residual per entry).
template <class T>
int layout_ops(const HostOperator *const *ops, int nops, DevOps &d, const std::vector<int> *hnet = nullptr,
               bool dense64 = false, bool acc64 = false, bool split = false) {
  const bool scaled = sizeof(T) == 4;  // FP32 storage keeps each block divided by its power-of-two scale
  dense64 = dense64 && scaled;
  split = split && dense64;
  d.acc64 = acc64 && scaled;
  d.split = split;
  const int n = ops[0]->n;
  for (int o = 1; o < nops; o++)
    if (ops[o]->perm != ops[0]->perm) return DFN_ERR_UNSUPPORTED;  // T and N share the cluster tree
  d.n = n;
  d.nops = nops;
  d.fp32 = sizeof(T) == 4;
  // w = [x_h | kXPad zeros | z of op 0 | z of op 1]
  std::vector<int64_t> zbase(nops + 1, n + kXPad);
  std::vector<std::vector<int>> zoff(nops);
  for (int o = 0; o < nops; o++) {
    const HostOperator &H = *ops[o];
    zoff[o].resize(H.l_k.size());
    int64_t z = 0;
    for (size_t b = 0; b < H.l_k.size(); b++) { zoff[o][b] = (int)z; z += H.l_k[b]; }
    zbase[o + 1] = zbase[o] + z;
  }
  const int64_t wsize = zbase[nops];
  // phase 1 tasks and V data
  std::vector<VTask> vt;
  std::vector<ZRed> zr;
  int zp = 0;
  int64_t S_total = 0, S_panel = 0;
  // V of each phase-1 task (a chunk of at most kVChunkRows rows of one block) is
  // stored strip by strip: a strip is the 64 x kVec rows one wave instruction covers,
  // and holds the strip's rows of column 0, then of column 1, ... (the last strip's
  // height padded with zero rows to a whole 16-byte vector).  The k loads of one
  // wave step are then one contiguous run of memory, and a task reads its data
  // front to back.
  constexpr int kVec = 16 / (int)sizeof(T);
  constexpr int kStrip = 64 * kVec;
  static_assert(kVChunkRows % kStrip == 0, "phase-1 chunks hold whole strips");
  // planned here, written by parallel_for below
  struct VFill { const double *src; int nb, k, r0, len; double vs; int64_t voff; };
  std::vector<VFill> vfill;
  int64_t vsize = 0;
  for (int o = 0; o < nops; o++) {
    const HostOperator &H = *ops[o];
    std::vector<double> vsc(H.l_k.size(), 1.0);
    if (scaled)
      parallel_for((int64_t)H.l_k.size(), [&](int64_t b) {
        vsc[b] = pow2_scale(H.l_v.data() + H.l_voff[b], (size_t)H.l_n[b] * H.l_k[b]);
      });
    for (size_t b = 0; b < H.l_k.size(); b++) {
      const int k = H.l_k[b], nb = H.l_n[b];
      if (k > kMaxRank) return DFN_ERR_UNSUPPORTED;
      const double *vb = H.l_v.data() + H.l_voff[b];
      const double vs = vsc[b];
      const int wz = (int)(zbase[o] + zoff[o][b]);
      const bool single = nb <= kVChunkRows;  // one chunk: z goes straight into w
      ZRed r{zp, 0, k, wz};
      for (int r0 = 0; r0 < nb; r0 += kVChunkRows) {
        const int len = std::min(kVChunkRows, nb - r0);
        const int64_t voff = vsize;
        for (int s0 = 0; s0 < len; s0 += kStrip) vsize += (int64_t)((std::min(kStrip, len - s0) + kVec - 1) / kVec * kVec) * k;
        vfill.push_back({vb, nb, k, r0, len, vs, voff});
        VTask t{};
        t.voff = voff;
        t.row0 = r0;
        t.len = len;
        t.k = k;
        t.xcol = H.l_c0[b];
        t.zp = single ? -1 : zp;
        t.scale = vs;
        t.wout = single ? wz : -1;
        t.net = hnet ? (*hnet)[H.l_c0[b]] : 0;
        vt.push_back(t);
        if (!single) zp += k;
        r.ntask++;
      }
      if (!single) zr.push_back(r);
    }
    S_total += (int64_t)H.l_v.size() + (int64_t)H.l_u.size() + (int64_t)H.d_data.size();
    S_panel += (int64_t)H.l_u.size() + (int64_t)H.d_data.size();
  }
  std::unique_ptr<T[]> vdata(new T[std::max<int64_t>(vsize, 1)]);
  parallel_for((int64_t)vfill.size(), [&](int64_t q) {
    const VFill &f = vfill[q];
    T *dst = vdata.get() + f.voff;
    for (int s0 = 0; s0 < f.len; s0 += kStrip) {
      const int h = std::min(kStrip, f.len - s0), hp = (h + kVec - 1) / kVec * kVec;
      for (int j = 0; j < f.k; j++) {
        const double *col = f.src + (size_t)j * f.nb + f.r0 + s0;
        for (int i = 0; i < hp; i++) dst[(size_t)j * hp + i] = i < h ? (T)(col[i] / f.vs) : (T)0;
      }
      dst += (size_t)hp * f.k;
    }
  });
  // phase 2 panels: per operator and leaf row cluster, all covering blocks
  const HostOperator &H0 = *ops[0];
  const int nleaf = (int)H0.leaf_r0.size();
  std::vector<int> row_leaf(n);
  for (int l = 0; l < nleaf; l++)
    for (int i = 0; i < H0.leaf_m[l]; i++) row_leaf[H0.leaf_r0[l] + i] = l;
  std::vector<std::vector<double>> dscale(nops), uscale(nops);
  int64_t S_dense = 0;
  for (int o = 0; o < nops; o++) {
    const HostOperator &H = *ops[o];
    S_dense += (int64_t)H.d_data.size();
    dscale[o].assign(H.d_m.size(), 1.0);
    uscale[o].assign(H.l_m.size(), 1.0);
    if (scaled) {
      if (!dense64)
        parallel_for((int64_t)H.d_m.size(), [&](int64_t b) {
          dscale[o][b] = pow2_scale(H.d_data.data() + H.d_off[b], (size_t)H.d_m[b] * H.d_n[b]);
        });
      parallel_for((int64_t)H.l_m.size(), [&](int64_t b) {
        uscale[o][b] = pow2_scale(H.l_u.data() + H.l_uoff[b], (size_t)H.l_m[b] * H.l_k[b]);
      });
    }
  }
  // columns of every panel: (w index, source pointer into the block, block scale);
  // panels are planned here and written by parallel_for below
  struct Col { int64_t wi; const double *src; double scale; };
  struct PFill { int64_t poff, col0; int L, Lp, m; };
  struct PanelSet {
    int esize;    // bytes per stored entry
    bool scales;  // per-column scales: FP32 storage's block scales, or the split form's task residual scale
    int pad;      // rows padded to a multiple of this many entries
    int maxc;     // columns per task at most (the w entries its kernel stages in LDS)
    std::vector<PTask> pt;
    std::vector<int> widx;
    std::vector<float> cscale;
    std::vector<PFill> pfill;
    int64_t psize = 0, entries = 0;
  };
  PanelSet setA{split ? kSplitBytes : dense64 ? 8 : (int)sizeof(T), scaled && (!dense64 || split),
                split ? 64 : dense64 ? 16 : 128 / (int)sizeof(T), split ? kSplitMaxCols : kPanelMaxCols},
      setB{(int)sizeof(T), scaled, 128 / (int)sizeof(T), kPanelMaxCols};
  std::vector<Col> cols;
  std::vector<std::vector<int>> fin((size_t)n * nops);
  int yp = 0;
  // tasks of the columns cols[c0, c0 + ncols) of one panel (rows r0.., m of them) into
  // set S; the first nd of them are near-field columns (compensated in FP32 panels)
  auto emit = [&](PanelSet &S, int o, int r0, int m, int net, int64_t c0, int64_t ncols, int64_t nd) {
    // rows padded with zero entries (on column 0 of w) to a multiple of 128 bytes, so
    // every row starts on a 128-byte line: a line shared by two rows is fetched
    // twice by the streaming loads (9 % more bytes fetched and 5-10 % more time
    // with 16-byte rows; the split set's rows are 64 entries: 256 bytes of FP32
    // values, and residual rows that fill a line per two rows); a zero entry adds an
    // exact zero to its lane's sum
    const int pad = S.pad;
    S.entries += ncols * m;
    for (int64_t q0 = 0; q0 < ncols; q0 += S.maxc) {
      const int L = (int)std::min<int64_t>(S.maxc, ncols - q0);
      const int Lp = (L + pad - 1) / pad * pad;
      PTask t{};
      t.poff = S.psize;
      t.widx = (int64_t)S.widx.size();
      t.m = m;
      t.L = Lp;
      t.yp = yp;
      t.nd = (int)std::max<int64_t>(0, std::min<int64_t>(L, nd - q0));
      t.net = net;
      for (int j = 0; j < Lp; j++) {
        S.widx.push_back(j < L ? (int)cols[c0 + q0 + j].wi : 0);
        if (S.scales) S.cscale.push_back(j < L ? (float)cols[c0 + q0 + j].scale : 1.0f);
      }
      S.pfill.push_back({S.psize, c0 + q0, L, Lp, m});
      S.psize += (int64_t)m * Lp;
      for (int r = 0; r < m; r++) fin[(size_t)o * n + r0 + r].push_back(yp + r);
      yp += m;
      S.pt.push_back(t);
    }
  };
  for (int o = 0; o < nops; o++) {
    const HostOperator &H = *ops[o];
    // covering lists: (kind, block) per leaf; dense first then low-rank, block order
    std::vector<std::vector<std::pair<int, int>>> cover(nleaf);
    for (size_t b = 0; b < H.d_m.size(); b++)
      for (int l = row_leaf[H.d_r0[b]]; l < nleaf && H0.leaf_r0[l] < H.d_r0[b] + H.d_m[b]; l++)
        cover[l].push_back({0, (int)b});
    for (size_t b = 0; b < H.l_m.size(); b++)
      for (int l = row_leaf[H.l_r0[b]]; l < nleaf && H0.leaf_r0[l] < H.l_r0[b] + H.l_m[b]; l++)
        cover[l].push_back({1, (int)b});
    for (int l = 0; l < nleaf; l++) {
      const int r0 = H0.leaf_r0[l], m = H0.leaf_m[l];
      // this panel's columns, near-field (dense) columns first
      const int64_t cbase = (int64_t)cols.size();
      int64_t ndense = 0;
      for (auto [kind, b] : cover[l]) {
        if (kind == 0) {
          const double *F = H.d_data.data() + H.d_off[b] + (r0 - H.d_r0[b]);
          for (int j = 0; j < H.d_n[b]; j++)
            cols.push_back({(int64_t)H.d_c0[b] + j, F + (int64_t)j * H.d_m[b], dscale[o][b]});
          ndense += H.d_n[b];
        } else {
          const double *U = H.l_u.data() + H.l_uoff[b] + (r0 - H.l_r0[b]);
          for (int j = 0; j < H.l_k[b]; j++)
            cols.push_back({zbase[o] + zoff[o][b] + j, U + (int64_t)j * H.l_m[b], uscale[o][b]});
        }
      }
      const int64_t ncols = (int64_t)cols.size() - cbase;
      const int net = hnet ? (*hnet)[r0] : 0;
      if (dense64) {  // near-field columns in FP64 (set A), then the U columns in FP32 (set B)
        emit(setA, o, r0, m, net, cbase, ndense, ndense);
        emit(setB, o, r0, m, net, cbase + ndense, ncols - ndense, 0);
      } else {
        emit(setA, o, r0, m, net, cbase, ncols, ndense);
      }
    }
  }
  // panel rows, zero entries padding each row to its 128-byte multiple
  auto fill = [&](const PanelSet &S, auto *pdata) {
    typedef std::remove_pointer_t<decltype(pdata)> E;
    parallel_for((int64_t)S.pfill.size(), [&](int64_t q) {
      const PFill &f = S.pfill[q];
      const Col *c = cols.data() + f.col0;
      for (int r = 0; r < f.m; r++) {
        E *row = pdata + f.poff + (int64_t)r * f.Lp;
        for (int j = 0; j < f.Lp; j++) row[j] = j < f.L ? (E)(c[j].src[r] / c[j].scale) : (E)0;
      }
    });
  };
  if (split) {  // near-field entries as hi = (float)entry and an 8-bit residual, two streams
    std::unique_ptr<float[]> ph(new float[std::max<int64_t>(setA.psize, 1)]);
    std::unique_ptr<signed char[]> pq(new signed char[std::max<int64_t>(setA.psize, 1)]);
    parallel_for((int64_t)setA.pfill.size(), [&](int64_t q) {
      const PFill &f = setA.pfill[q];  // built beside setA.pt[q]
      const Col *c = cols.data() + f.col0;
      // s: the task's largest |entry - (float)entry| over 127, so that every residual
      // rounds to an integer in [-127, 127]
      double rmax = 0.0;
      for (int j = 0; j < f.L; j++)
        for (int r = 0; r < f.m; r++) rmax = std::fmax(rmax, std::fabs(c[j].src[r] - (double)(float)c[j].src[r]));
      const float s = rmax > 0.0 ? (float)(rmax / 127.0) : 1.0f;
      std::fill_n(setA.cscale.begin() + setA.pt[q].widx, f.Lp, s);
      for (int r = 0; r < f.m; r++) {
        float *hr = ph.get() + f.poff + (int64_t)r * f.Lp;
        signed char *qr = pq.get() + f.poff + (int64_t)r * f.Lp;
        for (int j = 0; j < f.Lp; j++) {
          const double v = j < f.L ? c[j].src[r] : 0.0;
          const float h = (float)v;
          hr[j] = h;
          const double qd = std::nearbyint((v - (double)h) / (double)s);
          qr[j] = (signed char)std::max(-127.0, std::min(127.0, qd));
        }
      }
    });
    std::unique_ptr<T[]> pb(new T[std::max<int64_t>(setB.psize, 1)]);
    fill(setB, pb.get());
    DFN_TRY(d.pdata.upload(ph.get(), (size_t)setA.psize));
    DFN_TRY(d.pdata_lo.upload(pq.get(), (size_t)setA.psize));
    DFN_TRY(d.pdata_b.upload(pb.get(), (size_t)setB.psize));
  } else if (dense64) {
    std::unique_ptr<double[]> pa(new double[std::max<int64_t>(setA.psize, 1)]);
    std::unique_ptr<T[]> pb(new T[std::max<int64_t>(setB.psize, 1)]);
    fill(setA, pa.get());
    fill(setB, pb.get());
    DFN_TRY(d.pdata.upload(pa.get(), (size_t)setA.psize));
    DFN_TRY(d.pdata_b.upload(pb.get(), (size_t)setB.psize));
  } else {
    std::unique_ptr<T[]> pa(new T[std::max<int64_t>(setA.psize, 1)]);
    fill(setA, pa.get());
    DFN_TRY(d.pdata.upload(pa.get(), (size_t)setA.psize));
  }
  const std::vector<PTask> &pt = setA.pt;
  std::vector<int> fin_first((size_t)n * nops + 1, 0), fin_idx;
  for (size_t g = 0; g < fin.size(); g++) {
    fin_first[g] = (int)fin_idx.size();
    fin_idx.insert(fin_idx.end(), fin[g].begin(), fin[g].end());
  }
  fin_first[fin.size()] = (int)fin_idx.size();
  if (setA.widx.size() + setB.widx.size() > (size_t)INT32_MAX || zp < 0 || yp < 0) return DFN_ERR_UNSUPPORTED;
  DFN_TRY(d.perm.upload(H0.perm));
  DFN_TRY(d.vtasks.upload(vt));
  DFN_TRY(d.vdata.upload(vdata.get(), (size_t)vsize));
  DFN_TRY(d.zred.upload(zr));
  DFN_TRY(d.zpart.alloc(sizeof(double) * std::max(zp, 1)));
  DFN_TRY(d.ptasks.upload(pt));
  DFN_TRY(d.widx.upload(setA.widx));
  DFN_TRY(d.cscale.upload(setA.cscale));
  d.pa_fp32 = scaled && !dense64;
  d.npt_b = (int)setB.pt.size();
  if (dense64) {
    DFN_TRY(d.ptasks_b.upload(setB.pt));
    DFN_TRY(d.widx_b.upload(setB.widx));
    DFN_TRY(d.cscale_b.upload(setB.cscale));
  }
  DFN_TRY(d.w.alloc(sizeof(double) * wsize));
  DFN_HIP(hipMemset(d.w.p, 0, d.w.bytes));  // the zeros after x_h stay: only x_h and z are written
  if (d.fp32) {
    DFN_TRY(d.xf.alloc(sizeof(float) * ((size_t)n + 4)));
    DFN_HIP(hipMemset(d.xf.p, 0, sizeof(float) * ((size_t)n + 4)));
  }
  DFN_TRY(d.ypart.alloc(sizeof(double) * std::max(yp, 1)));
  DFN_TRY(d.fin_first.upload(fin_first));
  DFN_TRY(d.fin_idx.upload(fin_idx));
  d.nvt = (int)vt.size();
  d.nzr = (int)zr.size();
  d.kmax = 1;
  for (const VTask &t : vt) d.kmax = std::max(d.kmax, (int)t.k);
  d.kc = vtx_rank_ceiling(d.kmax);
  d.npt = (int)pt.size();
  d.S_total = S_total;
  d.S_panel = S_panel;
  // bytes: factors + x read + y written, per operator (x shared in Mode II)
  const double dense_extra = dense64 ? ((split ? (double)kSplitBytes : 8.0) - (double)sizeof(T)) * (double)S_dense : 0.0;
  d.alg_bytes = (double)sizeof(T) * (double)S_total + dense_extra + 8.0 * n + 8.0 * n * nops;
  d.panel_bytes = (double)setA.esize * (double)setA.entries + (double)setB.esize * (double)setB.entries +
                  4.0 * (double)((split ? setA.pt.size() : setA.cscale.size()) + setB.cscale.size()) + 8.0 * (double)yp;
  // one launch for phase 1 and the near-field panels where its rank ceiling keeps it
  // at 128 VGPRs without spills (ranks up to 20)
  d.near_ok = dense64 && d.acc64 && !split && vtx_rank_ceiling(d.kmax) <= 20;
  if (d.near_ok) {
    int64_t S_V = 0, nz = 0;  // nz: z entries written straight to w (single-chunk blocks)
    for (int o = 0; o < nops; o++) {
      S_V += (int64_t)ops[o]->l_v.size();
      for (size_t b = 0; b < ops[o]->l_k.size(); b++)
        if (ops[o]->l_n[b] <= kVChunkRows) nz += ops[o]->l_k[b];
    }
    int64_t rows_a = 0;
    for (const PTask &t : setA.pt) rows_a += t.m;
    d.fused_bytes = 4.0 * (double)S_V + 8.0 * (double)setA.entries + 8.0 * n + 8.0 * (double)(zp + nz) +
                    8.0 * (double)rows_a;
  }
  d.fused = false;
  d.vnt = false;
  const char *at = std::getenv("DFN_AUTOTUNE");
  if (!(at && std::string(at) == "0")) {
    DFN_TRY(autotune(d, vt, pt, setB.pt));
  } else if (const char *pe = std::getenv("DFN_VTX_PAIRS"); pe && pe[0] == '1') {
    const std::vector<VItem> items = vtx_items(vt, d.fp32 ? 4 : 2);  // over the tasks as built
    DFN_TRY(d.vitems.upload(items));
    d.nvi = (int)items.size();
    d.vpair = true;
  }
  // DFN_VTX_NT=1 / 0 takes / refuses the non-temporal FP64 factor loads
  const char *nt = std::getenv("DFN_VTX_NT");
  if (nt && (nt[0] == '0' || nt[0] == '1')) d.vnt = !d.fp32 && nt[0] == '1';
  // DFN_FUSE_NEAR=1 / 0 takes / refuses the single launch of phase 1 and the
  // near-field panels whatever the timing (the results are the same either way)
  const char *fuse = std::getenv("DFN_FUSE_NEAR");
  if (fuse && d.near_ok) {
    d.fused = std::string(fuse) == "1";
    if (d.fused) d.kc = vtx_rank_ceiling(d.kmax);
  }
  return DFN_OK;
}

// step-size proposal network (mlp.hip): forward pass of n feature rows already on
// the device, normalised outputs into dY; the output's physical value
struct Mlp;
int mlp_forward_device(Mlp *m, const float *dX, int n, float *dY, hipStream_t s);
float mlp_output_scale(const Mlp *m, float y);
Mlp *mlp_of(dfn_mlp h);

}  // namespace dfn
