// Host engine and C ABI (include/dfn.h): device-resident state of one problem,
// the device layout of the stress-transfer operators, and the solver loop of
// src/solver.f90:59-198 driving the kernels of kernels.hip on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "dfn_device.hpp"
#include "dfn_host.hpp"
#include "engine_common.hpp"

namespace dfn {

// Scratch of dfn_kernel_noalloc: the per-tile partial sums of the dense direct
// kernels, direct_chunks(N) * N doubles per operator (two operators in Mode II).
// 16 MiB in every GPU the library runs on, held until the process ends and
// shared by every call on that GPU.  Host code never uses the array's name as an
// address: each device's copy is reached through the pointer hipGetSymbolAddress
// returned with that device current.
constexpr size_t kNoAllocDoubles = (size_t)1 << 21;
static __device__ __align__(256) double scratch[kNoAllocDoubles];
constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_dev[kNoAllocMaxDevices];

// This GPU's copy of `scratch`, resolved once per device with that device current.
static int noalloc_scratch(double **out) {
  int dev = 0;
  DFN_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) return DFN_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    DFN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return DFN_ERR_HIP;
    g_scratch_dev[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_dev[dev];
  return DFN_OK;
}

struct Engine {
  HostProblem hp;
  HostOperator hop[2];
  bool have_op[2] = {false, false};
  int precision = DFN_PREC_FP64;
  hipStream_t stream = nullptr;
  std::unique_ptr<DevOps> ops64, ops32;  // FP64 layout; the reduced-precision layout of kind ops32_kind
  int ops32_kind = -1;
  // problem data on the device
  DMem node, elem, nrm, ds, a, b, Dc, f0, V0, nld, sld, fake, perm_star, kmin, Snk, src_elem, src_Q, src_tb, src_te,
      cross, dpart;
  DMem perm_cur, perm_next, permx_cur, permx_next;
  // state and stage buffers
  DMem Y, Ynew, yscal, sT, sN, kbuf, ytA, ytB, P, Pdot, scratch3, scratch3b;
  DMem cg_work, cg_sys, cg_res, cg_records, xrow_start, xrow_ent;
  // multi-RHS stress evaluation (dfn_kernel_multi): vector batches of up to mr_R
  // (a multiple of 64), input and output vectors and the FP64 layout's work arrays
  DMem mr_X, mr_Y, mr_W, mr_zp, mr_yp, mr_Yh;
  int mr_R = 0;
  bool coop_ok = true;
  int span_entries = 0;  // most crossing entries in one span of kCgSpan rows (cooperative CG)
  unsigned cg_seq = 0;   // sequence number of the last cooperative CG launch
  // compute_P kernel choice: 0 = by size (one workgroup per system up to
  // kCgLdsMax elements, else cooperative, else the plain one-workgroup kernel);
  // 1/2/3 = prefer the LDS / cooperative / plain kernel.  Set from the
  // environment variable DFN_CG_PATH ("lds", "coop", "plain") at dfn_create, so
  // that every kernel can be checked on small problems.
  int cg_path = 0;
  // test hook (DFN_CG_LATE_PART, read at dfn_create): workgroup n - 1 of every
  // cooperative compute_P system reports a message that never arrived; 0 off
  int cg_late_part = 0;
  DMem status;
  StatusWord *h_status = nullptr;  // pinned
  int *h_cgres = nullptr;          // pinned [kSlots*2]
  // CG systems: slots 0-5 are compute_P(t_try) and the five stage abscissae of a
  // first attempt; slots 6g..6g+5 the same for t_try / 2^g (rung g of a ladder of up
  // to kLevels), solved beside them so that a failed t_try needs no new launch while
  // the ladder lasts (the reference halves t_try and solves again).  A failing solve
  // runs its 200 iterations whatever else shares the launch, so the lower rungs cost
  // no time as long as their workgroups are resident together.
  static constexpr int kLevels = 4;
  static constexpr int kSlots = 6 * kLevels;
  int slot_base = 0;        // slot set the current attempt's stages read
  int last_halvings = 0;    // halvings of t_try at the start of the last step
  int coop_capacity = -1;
  // rungs (at most want, at least 1) whose 6 systems each fit one launch together:
  // the one-workgroup kernels always, the cooperative kernel when all their
  // workgroups can be resident together
  int ladder_levels(int want) {
    want = std::max(1, std::min(want, kLevels));
    const int N = hp.N;
    if (N <= kCgLdsMax && cg_path <= 1) return want;
    if (!coop_ok || cg_path == 3) return want;  // plain kernel: one workgroup per system
    if (coop_capacity < 0) coop_capacity = cg_coop_capacity(span_entries);
    const int wps = (N + kCgSpan - 1) / kCgSpan;
    while (want > 1 && 6 * want * wps > coop_capacity) want--;
    return want;
  }
  int last_pd_slot = -1;
  double time = 0, t_try = 100.0;
  dfn_run_stats st{};
  // event monitor: per accepted step the time and each fault's max |V|, and the
  // slip integral (trapezoid rule) per element
  bool monitor = false;
  DMem slip, fault_of, fvmax;
  unsigned long long *h_fvmax = nullptr;  // pinned [nfault]
  std::vector<double> mon_t, mon_v;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // while kt_on, every stress evaluation records an event pair around its dominant
  // kernel (the phase-2 panel launch, or the direct kernel); the pairs are summed
  // once their launches are complete
  bool kt_on = false;
  std::vector<hipEvent_t> kt_ev;  // pairs: kt_ev[2 i], kt_ev[2 i + 1]
  int kt_used = 0;
  double kt_ms = 0.0;
  int64_t kt_count = 0;
  // the step's compute_P batch can run on a second stream (default priority)
  // beside the k1 stress evaluation, which does not depend on it
  hipStream_t cg_stream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;

  double *k(int i) const { return kbuf.as<double>() + (size_t)i * 3 * hp.N; }
  double *cg_ptr(int slot, int which) const {  // 12 vectors of N per slot
    return cg_work.as<double>() + ((size_t)slot * 12 + which) * hp.N;
  }
  double *Pt(int slot) const { return cg_ptr(slot, 10); }
  double *Pd(int slot) const { return cg_ptr(slot, 11); }

  ~Engine() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (h_status) (void)hipHostFree(h_status);
    if (h_cgres) (void)hipHostFree(h_cgres);
    if (h_fvmax) (void)hipHostFree(h_fvmax);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    for (hipEvent_t ev : kt_ev) (void)hipEventDestroy(ev);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (cg_stream) {
      (void)hipStreamSynchronize(cg_stream);
      (void)hipStreamDestroy(cg_stream);
    }
    ops64.reset();
    ops32.reset();
    if (stream) (void)hipStreamDestroy(stream);
  }

  int init() {
    if (const char *cp = std::getenv("DFN_CG_PATH")) {
      const std::string v(cp);
      cg_path = v == "lds" ? 1 : v == "coop" ? 2 : v == "plain" ? 3 : 0;
    }
    if (const char *lp = std::getenv("DFN_CG_LATE_PART")) cg_late_part = std::max(0, std::atoi(lp));
    int dev = 0;
    DFN_HIP(hipGetDevice(&dev));
    DFN_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    DFN_HIP(hipEventCreate(&ev0));
    DFN_HIP(hipEventCreate(&ev1));
    DFN_HIP(hipStreamCreateWithFlags(&cg_stream, hipStreamNonBlocking));
    DFN_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    DFN_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    const int N = hp.N;
    DFN_TRY(node.upload(hp.node));
    DFN_TRY(elem.upload(hp.elem));
    DFN_TRY(nrm.upload(hp.nrm));
    DFN_TRY(ds.upload(hp.ds));
    DFN_TRY(a.upload(hp.a));
    DFN_TRY(b.upload(hp.b));
    DFN_TRY(Dc.upload(hp.Dc));
    DFN_TRY(f0.upload(hp.f0));
    DFN_TRY(V0.upload(hp.V0));
    DFN_TRY(nld.upload(hp.nld));
    DFN_TRY(sld.upload(hp.sld));
    DFN_TRY(fake.upload(hp.is_fake));
    DFN_TRY(perm_star.upload(hp.perm_star));
    DFN_TRY(kmin.upload(hp.kmin));
    DFN_TRY(Snk.upload(hp.Snk));
    DFN_TRY(src_elem.upload(hp.src_elem));
    DFN_TRY(src_Q.upload(hp.src_Q));
    DFN_TRY(src_tb.upload(hp.src_tb));
    DFN_TRY(src_te.upload(hp.src_te));
    DFN_TRY(cross.upload(hp.cross));
    if (hp.kernel == DFN_KERNEL_DIRECT) {
      DFN_TRY(dpart.alloc(sizeof(double) * (size_t)direct_chunks(N) * N * (hp.mode == DFN_MODE_II ? 2 : 1)));
      double *unused = nullptr;  // dfn_kernel_noalloc's scratch on this device
      DFN_TRY(noalloc_scratch(&unused));
    }
    // permeability_x (initialisation.f90:252-266): perm_x[k] = hmean(perm[k-1], perm[k]), k = 1..N-1
    std::vector<double> px(N + 1, 0.0);
    for (int kk = 1; kk < N; kk++)
      px[kk] = 2.0 * hp.perm0[kk - 1] * hp.perm0[kk] / (hp.perm0[kk - 1] + hp.perm0[kk]);
    DFN_TRY(perm_cur.upload(hp.perm0));
    DFN_TRY(perm_next.upload(hp.perm0));
    DFN_TRY(permx_cur.upload(px));
    DFN_TRY(permx_next.upload(px));
    const size_t n3 = 3 * (size_t)N * sizeof(double);
    DFN_TRY(Y.upload(hp.Y0));
    DFN_TRY(Ynew.alloc(n3));
    DFN_TRY(yscal.alloc(n3));
    DFN_TRY(sT.alloc(sizeof(double) * N));
    DFN_TRY(sN.alloc(sizeof(double) * N));
    DFN_HIP(hipMemset(sN.p, 0, sizeof(double) * N));
    DFN_TRY(kbuf.alloc(6 * n3));
    DFN_TRY(ytA.alloc(n3));
    DFN_TRY(ytB.alloc(n3));
    DFN_TRY(scratch3.alloc(n3));
    DFN_TRY(scratch3b.alloc(n3));
    DFN_TRY(P.upload(hp.P0));
    DFN_TRY(Pdot.alloc(sizeof(double) * N));
    DFN_HIP(hipMemset(Pdot.p, 0, sizeof(double) * N));
    DFN_TRY(cg_work.alloc(sizeof(double) * (size_t)kSlots * 12 * N));
    DFN_TRY(cg_res.alloc(sizeof(int) * kSlots * 2));
    DFN_TRY(cg_sys.alloc(sizeof(CgSys) * kSlots));
    DFN_TRY(status.alloc(sizeof(StatusWord)));
    const int wps = (N + kCgSpan - 1) / kCgSpan;
    DFN_TRY(cg_records.alloc(sizeof(unsigned long long) * kCgRecordWords * (size_t)kSlots * wps));
    // the cooperative CG's message words start with epoch 0, which no launch uses
    DFN_HIP(hipMemset(cg_records.p, 0, sizeof(unsigned long long) * kCgRecordWords * (size_t)kSlots * wps));
    DFN_HIP(hipMemset(cg_work.p, 0, sizeof(double) * (size_t)kSlots * 12 * N));
    // crossing rows of the stencil, per element in crossing order
    {
      std::vector<std::vector<std::pair<int, int>>> rows(N);
      for (int c = 0; c < hp.ncross(); c++) {
        const int64_t n1 = hp.cross[6 * c + 2], n2 = hp.cross[6 * c + 3];
        const int64_t e[4] = {n1 - 2, n1 - 1, n2 - 2, n2 - 1};
        for (int z = 0; z < 4; z++) rows[e[z]].push_back({c, z});
      }
      std::vector<int> start(N + 1, 0), ent;
      for (int k = 0; k < N; k++) {
        start[k] = (int)(ent.size() / 2);
        for (auto [c, z] : rows[k]) { ent.push_back(c); ent.push_back(z); }
      }
      start[N] = (int)(ent.size() / 2);
      for (int s0 = 0; s0 < N; s0 += kCgSpan)
        span_entries = std::max(span_entries, start[std::min(s0 + kCgSpan, N)] - start[s0]);
      if (ent.empty()) ent.assign(2, 0);
      DFN_TRY(xrow_start.upload(start));
      DFN_TRY(xrow_ent.upload(ent));
    }
    DFN_HIP(hipHostMalloc((void **)&h_status, sizeof(StatusWord), hipHostMallocDefault));
    DFN_HIP(hipHostMalloc((void **)&h_cgres, sizeof(int) * kSlots * 2, hipHostMallocDefault));
    time = 0.0;
    t_try = 100.0;
    return DFN_OK;
  }

  double kernel_scale() const {
    const double pi = 4.0 * std::atan(1.0);
    if (hp.mode == DFN_MODE_III) return hp.mu / (2.0 * pi);
    return hp.mu / pi * (1.0 - hp.cs * hp.cs / (hp.cp * hp.cp));
  }

  int relayout() {
    ops64.reset();
    ops32.reset();
    mr_R = 0;
    const int nops = hp.mode == DFN_MODE_II ? 2 : 1;
    const HostOperator *list[2] = {&hop[0], &hop[1]};
    for (int o = 0; o < nops; o++)
      if (!have_op[o]) return DFN_OK;  // incomplete until every operator exists
    ops64.reset(new DevOps);
    DFN_TRY(layout_ops<double>(list, nops, *ops64));
    if (precision != DFN_PREC_FP64) DFN_TRY(layout_reduced());
    return DFN_OK;
  }

  // the FP32 layout (DFN_PREC_FP32), FP32 low-rank factors read into FP64 products
  // and sums beside FP64 near-field blocks (DFN_PREC_MIXED) or the FP32 layout read
  // into FP64 products and sums (DFN_PREC_FP32_STORE), for the current precision
  int layout_reduced() {
    ops32.reset();
    ops32_kind = -1;
    if (!ops64) return DFN_OK;  // no operator yet
    const int nops = hp.mode == DFN_MODE_II ? 2 : 1;
    const HostOperator *list[2] = {&hop[0], &hop[1]};
    ops32.reset(new DevOps);
    const bool split = precision == DFN_PREC_MIXED_SPLIT, mixed = precision == DFN_PREC_MIXED || split;
    DFN_TRY(layout_ops<float>(list, nops, *ops32, nullptr, mixed, precision == DFN_PREC_FP32_STORE || mixed, split));
    ops32_kind = precision;
    return DFN_OK;
  }

  const DevOps *active_ops() const {
    return (precision != DFN_PREC_FP64 && ops32 && ops32_kind == precision) ? ops32.get() : ops64.get();
  }

  // stress rates sT (and sN) from the slip-rate vector d_V (device, original order)
  // the event pair for the next timed launch (kt_on), or nullptr
  hipEvent_t *kt_next() {
    if (!kt_on) return nullptr;
    if (kt_used >= 1024 && kt_collect() != DFN_OK) return nullptr;
    while ((int)kt_ev.size() < 2 * (kt_used + 1)) {
      hipEvent_t ev = nullptr;
      if (hipEventCreate(&ev) != hipSuccess) return nullptr;
      kt_ev.push_back(ev);
    }
    return &kt_ev[2 * kt_used++];
  }

  // adds the durations of the recorded pairs (waits for the last one)
  int kt_collect() {
    if (kt_used == 0) return DFN_OK;
    DFN_HIP(hipEventSynchronize(kt_ev[2 * kt_used - 1]));
    for (int i = 0; i < kt_used; i++) {
      float t = 0.0f;
      DFN_HIP(hipEventElapsedTime(&t, kt_ev[2 * i], kt_ev[2 * i + 1]));
      kt_ms += t;
      kt_count++;
    }
    kt_used = 0;
    return DFN_OK;
  }

  int apply_kernel(const double *d_V, double *d_sT, double *d_sN) {
    const int N = hp.N;
    hipEvent_t *kt = kt_next();
    if (hp.kernel == DFN_KERNEL_DIRECT) {
      if (kt) DFN_HIP(hipEventRecord(kt[0], stream));
      if (hp.mode == DFN_MODE_III)
        launch_direct3(node.as<double>(), d_V, dpart.as<double>(), N, kernel_scale(), d_sT, stream);
      else
        launch_direct2(node.as<double>(), elem.as<double>(), nrm.as<double>(), d_V, dpart.as<double>(), N,
                       kernel_scale(), d_sT, d_sN, stream);
      if (kt) DFN_HIP(hipEventRecord(kt[1], stream));
      return DFN_OK;
    }
    const DevOps *d = active_ops();
    if (!d) return DFN_ERR_STATE;
    double *w = d->w.as<double>();
    launch_gather(d_V, d->perm.as<int>(), w, N, stream, d->xf.as<float>());
    launch_phases(*d, w, stream, d->vtx_waves, d->panel_rows, nullptr, kt);
    launch_finalize(d->fin_first.as<int>(), d->fin_idx.as<int>(), d->perm.as<int>(), d->ypart.as<double>(), N,
                    d->nops, kernel_scale(), d_sT, d->nops > 1 ? d_sN : d_sT, stream);
    return DFN_OK;
  }

  // work arrays for multi-RHS batches of R vectors (R a multiple of 64)
  int mrhs_reserve(int R) {
    if (R <= mr_R) return DFN_OK;
    const DevOps &d = *ops64;
    const size_t n = hp.N, r = R;
    DFN_TRY(mr_X.alloc(sizeof(double) * n * r));
    DFN_TRY(mr_Y.alloc(sizeof(double) * n * r * d.nops));
    DFN_TRY(mr_W.alloc(d.w.bytes * r));
    DFN_TRY(mr_zp.alloc(d.zpart.bytes * r));
    DFN_TRY(mr_yp.alloc(d.ypart.bytes * r));
    DFN_TRY(mr_Yh.alloc(sizeof(double) * n * r * d.nops));
    mr_R = R;
    return DFN_OK;
  }

  // largest multi-RHS batch: at most 1024 vectors and about 48 GiB of work arrays
  int mrhs_batch(int nrhs) const {
    const DevOps &d = *ops64;
    const double per = (double)d.w.bytes + d.zpart.bytes + d.ypart.bytes + 8.0 * hp.N * (1 + 2 * d.nops);
    int cap = (int)std::min(1024.0, std::floor(48.0 * (1 << 30) / per / 64) * 64);
    cap = std::max(cap, 64);
    return std::min(cap, (nrhs + 63) / 64 * 64);
  }

  // stress rates of nb vectors (device, vector after vector, original order) with the
  // FP64 layout into mr_Y (T vectors, then N vectors); ev: events around phase 2
  int apply_multi(const double *dX, int nb, hipEvent_t *ev = nullptr) {
    const DevOps &d = *ops64;
    const int R = (nb + 63) / 64 * 64;
    DFN_TRY(mrhs_reserve(R));
    double *Y = mr_Y.as<double>();
    launch_mrhs(d, dX, nb, R, mr_W.as<double>(), mr_zp.as<double>(), mr_yp.as<double>(), mr_Yh.as<double>(),
                kernel_scale(), Y, Y + (size_t)hp.N * nb, stream, ev);
    DFN_HIP(hipGetLastError());
    return DFN_OK;
  }

  // compute_P for the time steps dts[0..nsys) into CG slots slot0.. (no host sync)
  int cg_launch(const double *dts, int nsys, int slot0) { return cg_launch_on(dts, nsys, slot0, stream); }

  // compute_P for nsys time steps into CG slots slot0.., issued on stream cs
  int cg_launch_on(const double *dts, int nsys, int slot0, hipStream_t cs) {
    const int N = hp.N;
    std::vector<CgSys> sys(nsys);
    for (int s = 0; s < nsys; s++) {
      const int sl = slot0 + s;
      CgSys &c = sys[s];
      c.dt = dts[s];
      c.diag = cg_ptr(sl, 0); c.A1 = cg_ptr(sl, 1); c.A2 = cg_ptr(sl, 2); c.A3 = cg_ptr(sl, 3);
      c.b = cg_ptr(sl, 4); c.x = cg_ptr(sl, 5); c.r = cg_ptr(sl, 6); c.p = cg_ptr(sl, 7); c.Ap = cg_ptr(sl, 8);
      c.Pt = Pt(sl); c.Pd = Pd(sl);
      c.result = cg_res.as<int>() + 2 * sl;
    }
    DFN_HIP(hipMemcpyAsync(cg_sys.as<CgSys>() + slot0, sys.data(), sizeof(CgSys) * nsys, hipMemcpyHostToDevice,
                           cs));
    // permeability update (fluid_diffusion.f90:67-75); diag keeps the values before it
    const double *px_diag = permx_cur.as<double>(), *p_diag = perm_cur.as<double>();
    const double *px = px_diag, *pp = p_diag;
    if (hp.coupling == 1) {
      DFN_HIP(hipMemcpyAsync(perm_next.p, perm_cur.p, sizeof(double) * N, hipMemcpyDeviceToDevice, cs));
      PermArgs pa{N, 1, perm_next.as<double>(), permx_next.as<double>(), perm_star.as<double>(), kmin.as<double>(),
                  Y.as<double>() + 2 * N, P.as<double>(), Snk.as<double>()};
      launch_perm_update(pa, cs);
      px = permx_next.as<double>();
      pp = perm_next.as<double>();
    }
    CgArgs A{};
    A.n = N;
    A.sys = cg_sys.as<CgSys>() + slot0;
    A.perm_x = px;
    A.perm_x_diag = px_diag;
    A.perm = pp;
    A.perm_diag = p_diag;
    A.ds = ds.as<double>();
    A.P = P.as<double>();
    A.cfac = hp.por * hp.visc * (hp.fcomp + hp.rcomp);
    A.time = time;
    A.tol = hp.tol;
    A.por = hp.por;
    A.beta = hp.fcomp + hp.rcomp;
    A.nsrc = (int)hp.src_elem.size();
    A.src_elem = src_elem.as<int64_t>();
    A.src_Q = src_Q.as<double>();
    A.src_tb = src_tb.as<double>();
    A.src_te = src_te.as<double>();
    A.ncross = hp.ncross();
    A.cross = cross.as<int64_t>();
    A.late_part = cg_late_part;
    bool launched = false;
    if (N <= kCgLdsMax && cg_path <= 1) {
      CrossRows X{xrow_start.as<int>(), xrow_ent.as<int>()};
      launched = launch_cg_lds(A, X, nsys, cs);
    }
    if (!launched && coop_ok && cg_path != 3) {
      CrossRows X{xrow_start.as<int>(), xrow_ent.as<int>()};
      if (++cg_seq >= kCgMaxSeq) {  // epochs would repeat: clear every message word first
        const int wps = (N + kCgSpan - 1) / kCgSpan;
        DFN_HIP(hipMemsetAsync(cg_records.p, 0, sizeof(unsigned long long) * kCgRecordWords * (size_t)kSlots * wps,
                               cs));
        DFN_HIP(hipMemsetAsync(cg_work.p, 0, sizeof(double) * (size_t)kSlots * 12 * N, cs));
        cg_seq = 1;
      }
      if (coop_capacity < 0) coop_capacity = cg_coop_capacity(span_entries);
      launched = launch_cg_coop(A, X, nsys, span_entries, coop_capacity, cg_seq, cg_records.as<unsigned long long>(),
                                cs);
      if (!launched) coop_ok = false;  // grid too large to be co-resident: one workgroup per system
    }
    if (!launched) launch_cg(A, nsys, cs);
    if (hp.coupling == 1) {  // the updated permeability becomes the state (swap buffers)
      std::swap(perm_cur.p, perm_next.p);
      std::swap(permx_cur.p, permx_next.p);
    }
    st.cg_solves += nsys;
    last_pd_slot = slot0 + nsys - 1;
    return DFN_OK;
  }

  // reads the results of CG slots slot0..; a solve whose workgroups lost contact
  // (result 2) is an error of the device, not a reason to halve the step
  int cg_collect(int slot0, int nsys, int *any_fail) {
    DFN_HIP(hipMemcpyAsync(h_cgres + 2 * slot0, cg_res.as<int>() + 2 * slot0, sizeof(int) * 2 * nsys,
                           hipMemcpyDeviceToHost, stream));
    DFN_HIP(hipStreamSynchronize(stream));
    *any_fail = 0;
    for (int s = slot0; s < slot0 + nsys; s++) {
      st.cg_iterations += h_cgres[2 * s + 1];
      if (h_cgres[2 * s] == 2) {
        std::fprintf(stderr, "dfn: compute_P workgroups lost contact (a message wait gave up)\n");
        return DFN_ERR_CG_TIMEOUT;
      }
      if (h_cgres[2 * s]) *any_fail = 1;
    }
    return DFN_OK;
  }

  RhsParams rhs_params() {
    return RhsParams{a.as<double>(), b.as<double>(), Dc.as<double>(), f0.as<double>(), V0.as<double>(),
                     nld.as<double>(), sld.as<double>(), fake.as<uint8_t>(), hp.mu / (2.0 * hp.cs),
                     hp.mode == DFN_MODE_III ? 1 : 0};
  }

  // Newton-V form (an extension; DFN_FORM_NEWTON_V): Y holds [slip since the switch |
  // psi | 0]; the balance's reference stresses, the last solved slip rates and the
  // switch time define the stress at every evaluation
  int form = DFN_FORM_REFERENCE;
  DMem nv_tau, nv_sig, nv_Vg;
  double nv_t0 = 0.0;

  NvArgs nv_args(double x) {
    return NvArgs{nv_tau.as<double>(), nv_sig.as<double>(), nv_Vg.as<double>(), (time - nv_t0) + x, 1};
  }

  int nv_enter() {
    const int N = hp.N;
    if (!nv_tau.p) {
      DFN_TRY(nv_tau.alloc(sizeof(double) * N));
      DFN_TRY(nv_sig.alloc(sizeof(double) * N));
      DFN_TRY(nv_Vg.alloc(sizeof(double) * N));
    }
    launch_nv_enter(rhs_params(), Y.as<double>(), P.as<double>(), N, nv_tau.as<double>(), nv_sig.as<double>(),
                    nv_Vg.as<double>(), stream);
    nv_t0 = time;
    DFN_HIP(hipStreamSynchronize(stream));
    return DFN_OK;
  }

  // the reference-form state [V | psi | sigma_N] of the Newton-V state into out (3N)
  int nv_state(double *out) {
    DFN_TRY(apply_kernel(Y.as<double>(), sT.as<double>(), sN.as<double>()));
    launch_nv_state(rhs_params(), nv_args(0.0), Y.as<double>(), P.as<double>(), sT.as<double>(), sN.as<double>(),
                    hp.N, out, stream);
    DFN_HIP(hipStreamSynchronize(stream));
    return DFN_OK;
  }

  int set_formulation(int f) {
    if (f != DFN_FORM_REFERENCE && f != DFN_FORM_NEWTON_V) return DFN_ERR_ARG;
    if (f == form) return DFN_OK;
    if (monitor || !out_dir.empty()) return DFN_ERR_STATE;
    if (hp.kernel == DFN_KERNEL_HMATRIX && !active_ops()) return DFN_ERR_STATE;
    DFN_TRY(sync_pdot());
    if (f == DFN_FORM_NEWTON_V) {
      DFN_TRY(nv_enter());
    } else {
      DFN_TRY(nv_state(scratch3.as<double>()));
      DFN_HIP(hipMemcpy(Y.p, scratch3.p, sizeof(double) * 3 * (size_t)hp.N, hipMemcpyDeviceToDevice));
    }
    form = f;
    return DFN_OK;
  }

  StageArgs stage_args(const double *ytemp, int slot, double *ynext, double h) {
    const int N = hp.N;
    StageArgs s{};
    s.n = N;
    s.p = rhs_params();
    if (form == DFN_FORM_NEWTON_V) {
      // the abscissa of this evaluation: k1 at x = t_try (h), stage s at A_s h
      static const double Ai[5] = {0.2, 0.3, 0.6, 1.0, 0.875};
      s.nv = nv_args(slot == 0 ? h : Ai[slot - 1] * h);
    }
    s.y = Y.as<double>();
    s.ytemp = ytemp;
    s.sT = sT.as<double>();
    s.sN = hp.mode == DFN_MODE_II ? sN.as<double>() : nullptr;
    s.Pt = Pt(slot_base + slot);
    s.Pd = Pd(slot_base + slot);
    for (int i = 0; i < 6; i++) s.k[i] = k(i);
    s.ynext = ynext;
    s.yscal = yscal.as<double>();
    s.h = h;
    s.t_try = t_try;
    s.tol = hp.tol;
    s.status = status.as<StatusWord>();
    return s;
  }

  // friction_law(x, Y) for a device Y into device Ydot (uses CG slot 0)
  int friction_law_dev(double x, const double *dY, double *dYd) {
    slot_base = 0;
    DFN_TRY(apply_kernel(dY, sT.as<double>(), sN.as<double>()));
    DFN_TRY(cg_launch(&x, 1, 0));
    int fail = 0;
    DFN_TRY(cg_collect(0, 1, &fail));
    st.rhs_evaluations++;
    if (fail) return DFN_ERR_CG;
    StageArgs s = stage_args(dY, 0, nullptr, 0.0);
    double *k0 = s.k[0];
    s.k[0] = dYd;
    launch_stage(0, s, stream);
    s.k[0] = k0;
    DFN_HIP(hipStreamSynchronize(stream));
    return DFN_OK;
  }

  // rkqs (ode_bs.f90:125-174) on the device state Y with k1 = dydx already in k(0)
  // and yscal set; x is the step-start abscissa.  On success Y holds the new state.
  // stages_ready: CG slots 1..5 already hold compute_P(x + A_i*htry) (launched with the
  // step's t_try solve)
  // rkqs_dev's answer when its first attempt ran before compute_P(t_try) had been
  // read (t0_unread) and that solve failed: the attempt is dropped
  static constexpr int kT0Failed = -1;

  int rkqs_dev(double x, double htry, double eps, double *hdid, double *hnext, bool stages_ready = false,
               bool t0_unread = false) {
    const int N = hp.N, n3 = 3 * N;
    const double SAFETY = 0.9, PGROW = -0.2, PSHRNK = -0.25, ERRCON = (double)1.89e-4f;
    double h = htry, errmax = 0.0;
    launch_restart_stage2(Y.as<double>(), k(0), h, n3, ytA.as<double>(), stream);
    for (;;) {
      static const double Ai[5] = {0.2, 0.3, 0.6, 1.0, 0.875};
      double dts[5];
      for (int s = 0; s < 5; s++) dts[s] = x + Ai[s] * h;
      if (!stages_ready) {
        slot_base = 0;
        DFN_TRY(cg_launch(dts, 5, 1));
      }
      else st.cg_solves += 5;
      stages_ready = false;
      DFN_HIP(hipMemsetAsync(status.p, 0, sizeof(StatusWord), stream));
      double *in = ytA.as<double>(), *out = ytB.as<double>();
      for (int s = 2; s <= 6; s++) {
        DFN_TRY(apply_kernel(in, sT.as<double>(), sN.as<double>()));
        StageArgs a = stage_args(in, s - 1, s == 6 ? Ynew.as<double>() : out, h);
        launch_stage(s, a, stream);
        st.rhs_evaluations++;
        std::swap(in, out);
      }
      DFN_HIP(hipMemcpyAsync(h_status, status.p, sizeof(StatusWord), hipMemcpyDeviceToHost, stream));
      int fail = 0;
      if (t0_unread) {  // compute_P(t_try) is read with the stage solves, in one copy
        t0_unread = false;
        DFN_TRY(cg_collect(slot_base, 6, &fail));
        if (h_cgres[2 * slot_base]) return kT0Failed;
      } else {
        DFN_TRY(cg_collect(slot_base + 1, 5, &fail));
      }
      if (fail) return DFN_ERR_CG;
      double em;
      std::memcpy(&em, &h_status->errmax_bits, sizeof(double));
      errmax = em / eps;
      if (errmax <= 1.0) break;
      st.rejected_attempts++;
      const double htemp = SAFETY * h * std::pow(errmax, PSHRNK);
      h = std::copysign(std::fmax(std::fabs(htemp), 0.1 * std::fabs(h)), h);
      const double xnew = x + h;
      if (xnew == x) return DFN_ERR_UNDERFLOW;
      launch_restart_stage2(Y.as<double>(), k(0), h, n3, ytA.as<double>(), stream);
    }
    if (errmax > ERRCON) *hnext = SAFETY * h * std::pow(errmax, PGROW);
    else *hnext = 5.0 * h;
    *hdid = h;
    std::swap(Y.p, Ynew.p);
    return DFN_OK;
  }

  // one iteration of src/solver.f90:59-198
  int step() {
    const int N = hp.N;
    if (t_try > 10.0)  // :64-77
      for (size_t s = 0; s < hp.src_tb.size(); s++)
        if ((time < hp.src_tb[s]) && (time + t_try > hp.src_tb[s])) t_try = 10.0;
    int fail = 0;
    // compute_P(t_try) (:81-92).  Without permeability evolution every compute_P of the
    // step depends on dt alone, so the five stage solves of a first attempt with
    // h = t_try (x + A_i h, x = 0) run in the same launch.
    const bool batch = hp.coupling != 1;
    // the solves for t_try / 2, / 4, ... run in the same launch when they fit beside
    // the others: one rung more than the last step needed halvings (none when it
    // needed none: they would only add work)
    const int levels = batch ? ladder_levels(1 + last_halvings) : 1;
    const int64_t halvings0 = st.cg_halvings;
    auto launch_try = [&](hipStream_t cs) -> int {
      if (!batch) return cg_launch_on(&t_try, 1, 0, cs);
      double dts[kSlots];
      double h = t_try;
      for (int g = 0; g < levels; g++, h = h / 2.0) {  // the values the halving below gives
        const double row[6] = {h, 0.0 + 0.2 * h, 0.0 + 0.3 * h, 0.0 + 0.6 * h, 0.0 + 1.0 * h, 0.0 + 0.875 * h};
        for (int s = 0; s < 6; s++) dts[6 * g + s] = row[s];
      }
      return cg_launch_on(dts, 6 * levels, 0, cs);
    };
    // The first batch of pressure solves runs on the side stream beside the k1
    // stress evaluation, which needs only Y (the cooperative CG too: its launch
    // keeps its workgroups co-resident, and the latency-bound solve then hides
    // behind the bandwidth-bound matvec: 4.31 -> 4.21 ms per step at 80k).
    DFN_HIP(hipEventRecord(ev_fork, stream));
    DFN_HIP(hipStreamWaitEvent(cg_stream, ev_fork, 0));
    DFN_TRY(launch_try(cg_stream));
    DFN_HIP(hipEventRecord(ev_join, cg_stream));
    DFN_TRY(apply_kernel(Y.as<double>(), sT.as<double>(), sN.as<double>()));
    DFN_HIP(hipStreamWaitEvent(stream, ev_join, 0));
    slot_base = 0;
    double t_did = 0, t_next = 0;
    // After a step that did not halve, stage 1 and the first attempt are issued
    // before compute_P(t_try)'s result is read (one host synchronisation less per
    // step); if that solve failed, the attempt's results are dropped and the step
    // halves as below, as the reference does before its k1 evaluation.
    const bool optimistic = batch && levels == 1 && last_halvings == 0;
    bool attempted = false;
    if (optimistic) {
      StageArgs s1 = stage_args(Y.as<double>(), 0, ytA.as<double>(), t_try);
      launch_stage(1, s1, stream);
      st.rhs_evaluations++;
      st.cg_solves -= 5;  // as below: the stages count their solves when they use them
      const int rc = rkqs_dev(0.0, t_try, hp.tol, &t_did, &t_next, true, true);
      if (rc == DFN_OK) attempted = true;
      else if (rc == kT0Failed) fail = 1;
      else return rc;
    } else {
      DFN_TRY(cg_collect(0, 1, &fail));
    }
    int rung = 0;  // the ladder rung whose solves belong to the current t_try
    while (fail) {  // :86-90: halve t_try until compute_P converges
      t_try = t_try / 2.0;
      st.cg_halvings++;
      if (rung + 1 < levels) {  // compute_P(t_try) already ran on the next rung
        rung++;
        slot_base = 6 * rung;
        DFN_TRY(cg_collect(slot_base, 1, &fail));
        continue;
      }
      rung = 0;  // a new ladder from this t_try
      slot_base = 0;
      DFN_TRY(launch_try(stream));
      DFN_TRY(cg_collect(0, 1, &fail));
    }
    last_halvings = (int)(st.cg_halvings - halvings0);
    if (!attempted) {
      if (batch) st.cg_solves -= 6 * levels - 1;  // the stages' solves are counted when used
      // k1 = f(t_try, Y) and yscal (:101-102): the stress rates are those evaluated
      // above (again if a dropped attempt overwrote them); compute_P(t_try) of the
      // RHS equals the call above unless the permeability evolves between calls
      if (optimistic) DFN_TRY(apply_kernel(Y.as<double>(), sT.as<double>(), sN.as<double>()));
      if (hp.coupling == 1) {
        DFN_TRY(cg_launch(&t_try, 1, 0));
        DFN_TRY(cg_collect(0, 1, &fail));
        if (fail) return DFN_ERR_CG;
      }
      StageArgs s1 = stage_args(Y.as<double>(), 0, ytA.as<double>(), t_try);
      launch_stage(1, s1, stream);
      st.rhs_evaluations++;
      DFN_TRY(rkqs_dev(0.0, t_try, hp.tol, &t_did, &t_next, batch));  // :111
    }
    t_try = t_next;  // :117-120
    if (t_try > 24 * 60 * 60) t_try = 24 * 60 * 60;
    // commit compute_P(t_did) (:174): the stage-5 solve used dt = 0 + 1.0*h = t_did
    if (hp.coupling == 1) {
      DFN_TRY(cg_launch(&t_did, 1, 5));
      DFN_TRY(cg_collect(5, 1, &fail));
    }
    if (!fail) {
      DFN_HIP(hipMemcpyAsync(P.p, Pt(slot_base + 4), sizeof(double) * N, hipMemcpyDeviceToDevice, stream));
      DFN_HIP(hipMemcpyAsync(Pdot.p, Pd(slot_base + 4), sizeof(double) * N, hipMemcpyDeviceToDevice, stream));
      if (hp.coupling == 1) {
        DFN_HIP(hipMemcpyAsync(P.p, Pt(5), sizeof(double) * N, hipMemcpyDeviceToDevice, stream));
        DFN_HIP(hipMemcpyAsync(Pdot.p, Pd(5), sizeof(double) * N, hipMemcpyDeviceToDevice, stream));
      }
    }
    last_pd_slot = -1;
    const double t_before = time;
    time = time + t_did;  // :177
    st.accepted_steps++;
    // cumtrapz integrates over the differences of the recorded times
    if (monitor) DFN_TRY(monitor_step(time - t_before));
    DFN_TRY(output_step());                         // save_data / write_data (:180-193)
    if (kt_on) DFN_TRY(kt_collect());
    if (h_status->tensile) return DFN_ERR_TENSILE;  // check_errors :412-420 on the accepted state
    return DFN_OK;
  }

  int monitor_start() {
    const int N = hp.N, nf = hp.nfault;
    std::vector<int> fo(N, -1);
    for (int f = 0; f < nf; f++)
      for (int64_t e = 0; e < hp.fnel[f]; e++) fo[hp.map1[f] - 1 + e] = f;
    DFN_TRY(fault_of.upload(fo));
    std::vector<double> z(N, 0.0);
    DFN_TRY(slip.upload(z));
    DFN_TRY(fvmax.alloc(sizeof(unsigned long long) * nf));
    if (!h_fvmax) DFN_HIP(hipHostMalloc((void **)&h_fvmax, sizeof(unsigned long long) * nf, hipHostMallocDefault));
    mon_t.clear();
    mon_v.clear();
    monitor = true;
    return DFN_OK;
  }

  // after an accepted step over the time difference h: Ynew holds the step-start
  // state, Y the new one
  // solver output (save_data / write_data, src/solver.f90:211-317)
  std::string out_dir;
  int64_t out_stride = 0, out_freq = 0, out_count = 0;
  std::vector<double> out_time;
  DMem out_buf;  // six (N x out_freq) fields, see launch_save_output

  int output_set(const char *dir, int64_t stride, int64_t freq) {
    out_dir.clear();
    out_count = 0;
    out_time.clear();
    if (!dir) return DFN_OK;
    if (stride < 1 || freq < 1) return DFN_ERR_ARG;
    const size_t bytes = sizeof(double) * 6 * (size_t)hp.N * (size_t)freq;
    if (out_buf.bytes < bytes) DFN_TRY(out_buf.alloc(bytes));
    out_dir = dir;
    out_stride = stride;
    out_freq = freq;
    return DFN_OK;
  }

  // after accepted step it_count = st.accepted_steps
  int output_step() {
    if (out_dir.empty()) return DFN_OK;
    const int64_t it_count = st.accepted_steps;
    if ((it_count - 1) % out_stride != 0) return DFN_OK;
    launch_save_output(Y.as<double>(), P.as<double>(), sT.as<double>(), perm_cur.as<double>(), out_buf.as<double>(),
                       hp.N, out_freq, out_count, stream);
    out_time.push_back(time);
    out_count++;
    if (out_count == out_freq) return output_write();
    return DFN_OK;
  }

  // dir/outputNNNNNNN with the out_count saves held, then none held
  int output_write() {
    if (out_dir.empty() || out_count == 0) return DFN_OK;
    const int64_t N = hp.N, F = out_freq, C = out_count;
    std::vector<double> h((size_t)(6 * N * F));
    DFN_HIP(hipMemcpyAsync(h.data(), out_buf.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, stream));
    DFN_HIP(hipStreamSynchronize(stream));
    std::vector<std::vector<double>> fld(6, std::vector<double>((size_t)(N * C)));
    for (int f = 0; f < 6; f++)
      for (int64_t e = 0; e < N; e++)
        std::memcpy(&fld[f][(size_t)(e * C)], &h[(size_t)(f * N * F + e * F)], sizeof(double) * (size_t)C);
    // theta = Dc / V0 * exp(psi) (:214)
    for (int64_t e = 0; e < N; e++)
      for (int64_t c = 0; c < C; c++) {
        double &v = fld[1][(size_t)(e * C + c)];
        v = hp.Dc[e] / hp.V0[e] * std::exp(v);
      }
    char name[32];
    std::snprintf(name, sizeof name, "/output%07lld", (long long)st.accepted_steps);
    const std::string path = out_dir + name;
    const char *names[7] = {"V", "theta", "P", "tractionTel", "tractionNel", "permeability", "time"};
    const int32_t ranks[7] = {2, 2, 2, 2, 2, 2, 1};
    int64_t dims[14];
    for (int i = 0; i < 6; i++) { dims[2 * i] = N; dims[2 * i + 1] = C; }
    dims[12] = C;
    dims[13] = 0;
    const double *data[7] = {fld[0].data(), fld[1].data(), fld[2].data(), fld[3].data(), fld[4].data(),
                             fld[5].data(), out_time.data()};
    const int rc = dfn_h5_write_f64(path.c_str(), 7, names, ranks, dims, data);
    out_count = 0;
    out_time.clear();
    return rc;
  }

  int monitor_step(double h) {
    const int nf = hp.nfault;
    launch_monitor(Ynew.as<double>(), Y.as<double>(), h, fault_of.as<int>(), hp.N, slip.as<double>(),
                   fvmax.as<unsigned long long>(), nf, stream);
    DFN_HIP(hipMemcpyAsync(h_fvmax, fvmax.p, sizeof(unsigned long long) * nf, hipMemcpyDeviceToHost, stream));
    DFN_HIP(hipStreamSynchronize(stream));
    mon_t.push_back(time);
    for (int f = 0; f < nf; f++) {
      double v;
      std::memcpy(&v, &h_fvmax[f], sizeof(double));
      mon_v.push_back(v);
    }
    return DFN_OK;
  }

  int sync_pdot() {
    if (last_pd_slot >= 0) {
      DFN_HIP(hipMemcpyAsync(Pdot.p, Pd(last_pd_slot), sizeof(double) * hp.N, hipMemcpyDeviceToDevice, stream));
      last_pd_slot = -1;
    }
    return DFN_OK;
  }
};

static std::string g_last_error;

}  // namespace dfn

using dfn::Engine;

struct dfn_handle_s {
  Engine e;
};

#define API extern "C" __attribute__((visibility("default")))
#include "api_guard.hpp"

static int device_ok() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return DFN_ERR_HIP;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return DFN_ERR_HIP;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    std::fprintf(stderr, "dfn: device is %s, this library is built for gfx950\n", prop.gcnArchName);
    return DFN_ERR_HIP;
  }
  return DFN_OK;
}

API int dfn_bench_stream(int64_t bytes, int32_t reps, double *gbps) {
  DFN_API_BEGIN
  if (bytes < 16 || reps < 1 || !gbps) return DFN_ERR_ARG;
  int st = device_ok();
  if (st) return st;
  bytes &= ~(int64_t)15;
  dfn::DMem a, b;
  DFN_TRY(a.alloc((size_t)bytes));
  DFN_TRY(b.alloc((size_t)bytes));
  hipStream_t s;
  hipEvent_t e0, e1;
  DFN_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  DFN_HIP(hipEventCreate(&e0));
  DFN_HIP(hipEventCreate(&e1));
  DFN_HIP(hipMemsetAsync(a.p, 0, (size_t)bytes, s));
  dfn::launch_stream_copy(a.p, b.p, bytes, s);
  DFN_HIP(hipEventRecord(e0, s));
  for (int r = 0; r < reps; r++) dfn::launch_stream_copy(a.p, b.p, bytes, s);
  DFN_HIP(hipEventRecord(e1, s));
  DFN_HIP(hipEventSynchronize(e1));
  float ms = 0;
  DFN_HIP(hipEventElapsedTime(&ms, e0, e1));
  *gbps = 2.0 * (double)bytes / (ms * 1e-3 / reps) / 1e9;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  return DFN_OK;
  DFN_API_END
}

API int dfn_device_info(char *buf, int32_t len) {
  DFN_API_BEGIN
  if (!buf || len < 1) return DFN_ERR_ARG;
  int n = 0;
  std::string s;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
    s = "no HIP device";
  } else {
    hipDeviceProp_t prop;
    (void)hipGetDeviceProperties(&prop, 0);
    char tmp[512];
    std::snprintf(tmp, sizeof(tmp), "%s %s CUs=%d mem=%.1fGB", prop.name, prop.gcnArchName, prop.multiProcessorCount,
                  prop.totalGlobalMem / 1e9);
    s = tmp;
  }
  std::strncpy(buf, s.c_str(), (size_t)len - 1);
  buf[len - 1] = 0;
  return DFN_OK;
  DFN_API_END
}

API int dfn_create(const dfn_problem *p, dfn_handle *out) {
  DFN_API_BEGIN
  if (!out) return DFN_ERR_ARG;
  *out = nullptr;
  int st = device_ok();
  if (st) return st;
  dfn_handle h = new (std::nothrow) dfn_handle_s;
  if (!h) return DFN_ERR_NOMEM;
  st = dfn::setup_problem(p, h->e.hp);
  if (st == DFN_OK) st = h->e.init();
  if (st != DFN_OK) { delete h; return st; }
  *out = h;
  return DFN_OK;
  DFN_API_END
}

API int dfn_destroy(dfn_handle h) {
  DFN_API_BEGIN
  if (!h) return DFN_ERR_ARG;
  delete h;
  return DFN_OK;
  DFN_API_END
}

API int dfn_build_hmat(dfn_handle h) {
  DFN_API_BEGIN
  if (!h) return DFN_ERR_ARG;
  Engine &e = h->e;
  if (e.hp.kernel != DFN_KERNEL_HMATRIX) return DFN_OK;
  // construction on the GPU (dense leaves and ACA) unless DFN_HBUILD=host asks
  // for the all-host path; both give the same operator
  const char *hb = std::getenv("DFN_HBUILD");
  const bool on_host = hb && std::string(hb) == "host";
  auto build = [&](int op) {
    return on_host ? dfn::build_operator(e.hp, op, e.hop[op]) : dfn::build_operator_device(e.hp, op, e.hop[op], e.stream);
  };
  if (e.hp.mode == DFN_MODE_II) {  // N then T (src/initialisation.f90:222-246)
    DFN_TRY(build(DFN_OP_N));
    e.have_op[DFN_OP_N] = true;
  }
  DFN_TRY(build(DFN_OP_T));
  e.have_op[DFN_OP_T] = true;
  const auto t0 = std::chrono::steady_clock::now();
  const int st = e.relayout();
  if (std::getenv("DFN_VERBOSE"))
    std::fprintf(stderr, "[dfn build] %-28s %8.3f s\n", "device layout + upload",
                 std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  return st;
  DFN_API_END
}

API int dfn_set_precision(dfn_handle h, int32_t prec) {
  DFN_API_BEGIN
  if (!h || (prec < DFN_PREC_FP64 || prec > DFN_PREC_MIXED_SPLIT)) return DFN_ERR_ARG;
  h->e.precision = prec;
  if (prec != DFN_PREC_FP64 && h->e.ops32_kind != prec) return h->e.layout_reduced();
  return DFN_OK;
  DFN_API_END
}

API int dfn_tuning(dfn_handle h, int32_t *choice, double *ms) {
  DFN_API_BEGIN
  if (!h || !choice || !ms) return DFN_ERR_ARG;
  const dfn::DevOps *d = h->e.active_ops();
  if (h->e.hp.kernel != DFN_KERNEL_HMATRIX || !d) return DFN_ERR_STATE;
  choice[0] = d->vtx_waves;
  choice[1] = d->vtx_order;
  choice[2] = d->panel_rows;
  choice[3] = d->panel_order;
  choice[4] = d->panel_split;
  choice[5] = d->fused ? 1 : 0;
  choice[6] = d->kc;
  choice[7] = d->vpair ? 1 : 0;
  ms[0] = d->tune_ms_default;
  ms[1] = d->tune_ms;
  return DFN_OK;
  DFN_API_END
}

API int dfn_hmat_stats(dfn_handle h, int32_t op, int64_t *stats) {
  DFN_API_BEGIN
  if (!h || !stats || op < 0 || op > 1) return DFN_ERR_ARG;
  const dfn::HostOperator &H = h->e.hop[op];
  if (!h->e.have_op[op]) return DFN_ERR_STATE;
  int64_t mr = 0, sr = 0;
  for (int k : H.l_k) { mr = std::max<int64_t>(mr, k); sr += k; }
  const int64_t s[10] = {(int64_t)H.d_m.size(), (int64_t)H.l_m.size(), (int64_t)H.d_data.size(),
                         (int64_t)H.l_u.size(), (int64_t)H.l_v.size(), mr, H.nbclus, H.nbblclus, H.n, sr};
  std::memcpy(stats, s, sizeof(s));
  return DFN_OK;
  DFN_API_END
}

API int dfn_hmat_export(dfn_handle h, int32_t op, int32_t *perm, int32_t *d_r0, int32_t *d_c0, int32_t *d_m,
                        int32_t *d_n, int64_t *d_off, double *d_data, int32_t *l_r0, int32_t *l_c0, int32_t *l_m,
                        int32_t *l_n, int32_t *l_k, int64_t *l_uoff, int64_t *l_voff, double *l_u, double *l_v) {
  DFN_API_BEGIN
  if (!h || op < 0 || op > 1) return DFN_ERR_ARG;
  if (!h->e.have_op[op]) return DFN_ERR_STATE;
  const dfn::HostOperator &H = h->e.hop[op];
  auto cp = [](auto *dst, const auto &v) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
  };
  cp(perm, H.perm); cp(d_r0, H.d_r0); cp(d_c0, H.d_c0); cp(d_m, H.d_m); cp(d_n, H.d_n); cp(d_off, H.d_off);
  cp(d_data, H.d_data); cp(l_r0, H.l_r0); cp(l_c0, H.l_c0); cp(l_m, H.l_m); cp(l_n, H.l_n); cp(l_k, H.l_k);
  cp(l_uoff, H.l_uoff); cp(l_voff, H.l_voff); cp(l_u, H.l_u); cp(l_v, H.l_v);
  return DFN_OK;
  DFN_API_END
}

API int dfn_hmat_import(dfn_handle h, int32_t op, int32_t n, const int32_t *perm, int32_t nd, const int32_t *d_r0,
                        const int32_t *d_c0, const int32_t *d_m, const int32_t *d_n, const int64_t *d_off,
                        const double *d_data, int32_t nl, const int32_t *l_r0, const int32_t *l_c0,
                        const int32_t *l_m, const int32_t *l_n, const int32_t *l_k, const int64_t *l_uoff,
                        const int64_t *l_voff, const double *l_u, const double *l_v) {
  DFN_API_BEGIN
  if (!h || op < 0 || op > 1 || !perm || n != h->e.hp.N || nd < 0 || nl < 0) return DFN_ERR_ARG;
  if (op == DFN_OP_N && h->e.hp.mode != DFN_MODE_II) return DFN_ERR_ARG;
  dfn::HostOperator H;
  H.n = n;
  H.perm.assign(perm, perm + n);
  std::vector<int> seen(n, 0);
  for (int v : H.perm) {
    if (v < 0 || v >= n || seen[v]) return DFN_ERR_ARG;
    seen[v] = 1;
  }
  H.d_r0.assign(d_r0, d_r0 + nd); H.d_c0.assign(d_c0, d_c0 + nd); H.d_m.assign(d_m, d_m + nd);
  H.d_n.assign(d_n, d_n + nd); H.d_off.assign(d_off, d_off + nd);
  H.l_r0.assign(l_r0, l_r0 + nl); H.l_c0.assign(l_c0, l_c0 + nl); H.l_m.assign(l_m, l_m + nl);
  H.l_n.assign(l_n, l_n + nl); H.l_k.assign(l_k, l_k + nl); H.l_uoff.assign(l_uoff, l_uoff + nl);
  H.l_voff.assign(l_voff, l_voff + nl);
  int64_t ds = 0, us = 0, vs = 0;
  std::vector<int> cut(n + 1, 0);  // row starts of all blocks -> leaf row clusters
  for (int b = 0; b < nd; b++) {
    if (d_m[b] < 1 || d_n[b] < 1 || d_r0[b] < 0 || d_c0[b] < 0 || d_r0[b] + d_m[b] > n || d_c0[b] + d_n[b] > n)
      return DFN_ERR_ARG;
    ds = std::max<int64_t>(ds, d_off[b] + (int64_t)d_m[b] * d_n[b]);
    cut[d_r0[b]] = 1; cut[d_r0[b] + d_m[b]] = 1;
  }
  for (int b = 0; b < nl; b++) {
    if (l_m[b] < 1 || l_n[b] < 1 || l_k[b] < 0 || l_k[b] > dfn::kMaxRank || l_r0[b] < 0 || l_c0[b] < 0 ||
        l_r0[b] + l_m[b] > n || l_c0[b] + l_n[b] > n)
      return DFN_ERR_ARG;
    us = std::max<int64_t>(us, l_uoff[b] + (int64_t)l_m[b] * l_k[b]);
    vs = std::max<int64_t>(vs, l_voff[b] + (int64_t)l_n[b] * l_k[b]);
    cut[l_r0[b]] = 1; cut[l_r0[b] + l_m[b]] = 1;
  }
  H.d_data.assign(d_data, d_data + ds);
  H.l_u.assign(l_u, l_u + us);
  H.l_v.assign(l_v, l_v + vs);
  cut[0] = 1; cut[n] = 1;
  int prev = 0;
  for (int i = 1; i <= n; i++)
    if (cut[i]) { H.leaf_r0.push_back(prev); H.leaf_m.push_back(i - prev); prev = i; }
  Engine &e = h->e;
  // in Mode II both operators must share row splits: merge with the other operator's
  const int other = 1 - op;
  if (e.hp.mode == DFN_MODE_II && e.have_op[other]) {
    std::vector<int> c2(n + 1, 0);
    for (size_t l = 0; l < H.leaf_r0.size(); l++) c2[H.leaf_r0[l]] = 1;
    for (size_t l = 0; l < e.hop[other].leaf_r0.size(); l++) c2[e.hop[other].leaf_r0[l]] = 1;
    c2[n] = 1;
    std::vector<int> r0s, ms;
    prev = 0;
    for (int i = 1; i <= n; i++)
      if (c2[i]) { r0s.push_back(prev); ms.push_back(i - prev); prev = i; }
    H.leaf_r0 = r0s; H.leaf_m = ms;
    e.hop[other].leaf_r0 = r0s; e.hop[other].leaf_m = ms;
  }
  e.hop[op] = std::move(H);
  e.have_op[op] = true;
  return e.relayout();
  DFN_API_END
}

API int dfn_set_state(dfn_handle h, const double *Yh, const double *Ph, double time) {
  DFN_API_BEGIN
  if (!h) return DFN_ERR_ARG;
  Engine &e = h->e;
  const size_t N = e.hp.N;
  if (e.form == DFN_FORM_NEWTON_V && Yh) {  // a reference-form state: leave and re-enter the Newton-V form
    DFN_TRY(e.set_formulation(DFN_FORM_REFERENCE));
    DFN_HIP(hipMemcpy(e.Y.p, Yh, 3 * N * sizeof(double), hipMemcpyHostToDevice));
    if (Ph) DFN_HIP(hipMemcpy(e.P.p, Ph, N * sizeof(double), hipMemcpyHostToDevice));
    e.time = time;
    return e.set_formulation(DFN_FORM_NEWTON_V);
  }
  if (Yh) DFN_HIP(hipMemcpy(e.Y.p, Yh, 3 * N * sizeof(double), hipMemcpyHostToDevice));
  if (Ph) DFN_HIP(hipMemcpy(e.P.p, Ph, N * sizeof(double), hipMemcpyHostToDevice));
  e.time = time;
  return DFN_OK;
  DFN_API_END
}

API int dfn_set_formulation(dfn_handle h, int32_t form) {
  DFN_API_BEGIN
  if (!h) return DFN_ERR_ARG;
  return h->e.set_formulation(form);
  DFN_API_END
}

API int dfn_get_state(dfn_handle h, double *Yh, double *Ph, double *Pdh, double *time) {
  DFN_API_BEGIN
  if (!h) return DFN_ERR_ARG;
  Engine &e = h->e;
  const size_t N = e.hp.N;
  DFN_TRY(e.sync_pdot());
  DFN_HIP(hipStreamSynchronize(e.stream));
  if (Yh && e.form == DFN_FORM_NEWTON_V) {
    DFN_TRY(e.nv_state(e.scratch3.as<double>()));
    DFN_HIP(hipMemcpy(Yh, e.scratch3.p, 3 * N * sizeof(double), hipMemcpyDeviceToHost));
  } else if (Yh) {
    DFN_HIP(hipMemcpy(Yh, e.Y.p, 3 * N * sizeof(double), hipMemcpyDeviceToHost));
  }
  if (Ph) DFN_HIP(hipMemcpy(Ph, e.P.p, N * sizeof(double), hipMemcpyDeviceToHost));
  if (Pdh) DFN_HIP(hipMemcpy(Pdh, e.Pdot.p, N * sizeof(double), hipMemcpyDeviceToHost));
  if (time) *time = e.time;
  return DFN_OK;
  DFN_API_END
}

API int dfn_get_stats(dfn_handle h, dfn_run_stats *s) {
  DFN_API_BEGIN
  if (!h || !s) return DFN_ERR_ARG;
  Engine &e = h->e;
  *s = e.st;
  s->time = e.time;
  s->t_try = e.t_try;
  std::vector<double> V(e.hp.N);
  DFN_HIP(hipStreamSynchronize(e.stream));
  const void *Vd = e.Y.p;
  if (e.form == DFN_FORM_NEWTON_V) {
    DFN_TRY(e.nv_state(e.scratch3.as<double>()));
    Vd = e.scratch3.p;
  }
  DFN_HIP(hipMemcpy(V.data(), Vd, V.size() * sizeof(double), hipMemcpyDeviceToHost));
  s->max_V = *std::max_element(V.begin(), V.end());
  return DFN_OK;
  DFN_API_END
}

API int dfn_kernel(dfn_handle h, const double *slip, double *sTh, double *sNh) {
  DFN_API_BEGIN
  if (!h || !slip || !sTh) return DFN_ERR_ARG;
  Engine &e = h->e;
  const size_t N = e.hp.N;
  if (e.hp.mode == DFN_MODE_II && !sNh) return DFN_ERR_ARG;
  double *dV = e.scratch3.as<double>();
  DFN_HIP(hipMemcpyAsync(dV, slip, N * sizeof(double), hipMemcpyHostToDevice, e.stream));
  DFN_TRY(e.apply_kernel(dV, e.sT.as<double>(), e.sN.as<double>()));
  DFN_HIP(hipMemcpyAsync(sTh, e.sT.p, N * sizeof(double), hipMemcpyDeviceToHost, e.stream));
  if (e.hp.mode == DFN_MODE_II)
    DFN_HIP(hipMemcpyAsync(sNh, e.sN.p, N * sizeof(double), hipMemcpyDeviceToHost, e.stream));
  DFN_HIP(hipStreamSynchronize(e.stream));
  return DFN_OK;
  DFN_API_END
}

API int dfn_kernel_multi(dfn_handle h, int32_t nrhs, const double *slip, double *sTh, double *sNh) {
  DFN_API_BEGIN
  if (!h || nrhs < 0 || (nrhs > 0 && (!slip || !sTh))) return DFN_ERR_ARG;
  Engine &e = h->e;
  const bool mode2 = e.hp.mode == DFN_MODE_II;
  if (mode2 && nrhs > 0 && !sNh) return DFN_ERR_ARG;
  if (e.hp.kernel != DFN_KERNEL_HMATRIX || !e.ops64) return DFN_ERR_STATE;
  const size_t N = e.hp.N;
  const int B = e.mrhs_batch(nrhs);
  for (int r0 = 0; r0 < nrhs; r0 += B) {
    const int nb = std::min(B, nrhs - r0);
    DFN_TRY(e.mrhs_reserve((nb + 63) / 64 * 64));
    DFN_HIP(hipMemcpyAsync(e.mr_X.p, slip + r0 * N, nb * N * sizeof(double), hipMemcpyHostToDevice, e.stream));
    DFN_TRY(e.apply_multi(e.mr_X.as<double>(), nb));
    const double *Y = e.mr_Y.as<double>();
    DFN_HIP(hipMemcpyAsync(sTh + r0 * N, Y, nb * N * sizeof(double), hipMemcpyDeviceToHost, e.stream));
    if (mode2)
      DFN_HIP(hipMemcpyAsync(sNh + r0 * N, Y + nb * N, nb * N * sizeof(double), hipMemcpyDeviceToHost, e.stream));
    DFN_HIP(hipStreamSynchronize(e.stream));
  }
  return DFN_OK;
  DFN_API_END
}

API int dfn_bench_kernel_multi(dfn_handle h, int32_t nrhs, const double *slip, int32_t reps, double *out) {
  DFN_API_BEGIN
  if (!h || !slip || !out || nrhs < 1 || reps < 1) return DFN_ERR_ARG;
  Engine &e = h->e;
  if (e.hp.kernel != DFN_KERNEL_HMATRIX || !e.ops64) return DFN_ERR_STATE;
  if (e.mrhs_batch(nrhs) < nrhs) return DFN_ERR_UNSUPPORTED;
  const size_t N = e.hp.N;
  DFN_TRY(e.mrhs_reserve((nrhs + 63) / 64 * 64));
  DFN_HIP(hipMemcpyAsync(e.mr_X.p, slip, nrhs * N * sizeof(double), hipMemcpyHostToDevice, e.stream));
  DFN_TRY(e.apply_multi(e.mr_X.as<double>(), nrhs));  // warm
  struct Events {
    hipEvent_t e[2] = {nullptr, nullptr};
    ~Events() {
      for (hipEvent_t x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } evs;
  hipEvent_t *ev = evs.e;
  DFN_HIP(hipEventCreate(&ev[0]));
  DFN_HIP(hipEventCreate(&ev[1]));
  float t = 0, tp = 0, tsum = 0;
  DFN_HIP(hipStreamSynchronize(e.stream));
  DFN_HIP(hipEventRecord(e.ev0, e.stream));
  for (int r = 0; r < reps; r++) DFN_TRY(e.apply_multi(e.mr_X.as<double>(), nrhs));
  DFN_HIP(hipEventRecord(e.ev1, e.stream));
  DFN_HIP(hipEventSynchronize(e.ev1));
  DFN_HIP(hipEventElapsedTime(&t, e.ev0, e.ev1));
  for (int r = 0; r < reps; r++) {  // the phase-2 kernel alone, launch by launch
    DFN_TRY(e.apply_multi(e.mr_X.as<double>(), nrhs, ev));
    DFN_HIP(hipEventSynchronize(ev[1]));
    DFN_HIP(hipEventElapsedTime(&tp, ev[0], ev[1]));
    tsum += tp;
  }
  const dfn::DevOps &d = *e.ops64;
  const int R = (nrhs + 63) / 64 * 64;
  out[0] = (double)t / reps;
  out[1] = (double)tsum / reps;
  out[2] = 2.0 * (double)d.S_panel * R;                           // phase-2 flops of one launch
  out[3] = 8.0 * (double)d.S_total + 16.0 * (double)N * nrhs * d.nops;  // matrix once + vectors in/out
  return DFN_OK;
  DFN_API_END
}

API int dfn_kernel_noalloc(dfn_handle h, const double *slip, double *sT, double *sN, void *stream) {
  DFN_API_BEGIN
  if (!h || !slip || !sT) return DFN_ERR_ARG;
  Engine &e = h->e;
  if (e.hp.kernel != DFN_KERNEL_DIRECT) return DFN_ERR_UNSUPPORTED;
  const int N = e.hp.N;
  const bool mode2 = e.hp.mode == DFN_MODE_II;
  if (mode2 && !sN) return DFN_ERR_ARG;
  const size_t need = (size_t)dfn::direct_chunks(N) * (size_t)N * (mode2 ? 2 : 1);
  if (need > dfn::kNoAllocDoubles) {
    std::fprintf(stderr,
                 "dfn_kernel_noalloc: N = %d needs %zu scratch doubles, more than the %zu it holds per GPU "
                 "(use dfn_kernel)\n",
                 N, need, dfn::kNoAllocDoubles);
    return DFN_ERR_ARG;
  }
  double *part = nullptr;
  DFN_TRY(dfn::noalloc_scratch(&part));
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (mode2)
    dfn::launch_direct2(e.node.as<double>(), e.elem.as<double>(), e.nrm.as<double>(), slip, part, N,
                        e.kernel_scale(), sT, sN, s);
  else
    dfn::launch_direct3(e.node.as<double>(), slip, part, N, e.kernel_scale(), sT, s);
  DFN_HIP(hipGetLastError());
  return DFN_OK;
  DFN_API_END
}

API int dfn_hgemv(dfn_handle h, int32_t op, double alpha, double beta, const double *x, double *y) {
  DFN_API_BEGIN
  if (!h || !x || !y || op < 0 || op > 1) return DFN_ERR_ARG;
  Engine &e = h->e;
  if (e.hp.kernel != DFN_KERNEL_HMATRIX || !e.active_ops()) return DFN_ERR_STATE;
  const dfn::DevOps *d = e.active_ops();
  if (op >= d->nops) return DFN_ERR_ARG;
  const int N = e.hp.N;
  // x is in H order: apply with the identity scatter by undoing the permutation on the host
  const std::vector<int> &perm = e.hop[0].perm;
  std::vector<double> xo(N), yT(N), yN(N);
  for (int i = 0; i < N; i++) xo[perm[i]] = x[i];
  double *dV = e.scratch3.as<double>();
  double *dT = e.scratch3b.as<double>(), *dN = dT + N;
  DFN_HIP(hipMemcpyAsync(dV, xo.data(), N * sizeof(double), hipMemcpyHostToDevice, e.stream));
  double *w = d->w.as<double>();
  dfn::launch_gather(dV, d->perm.as<int>(), w, N, e.stream, d->xf.as<float>());
  dfn::launch_phases(*d, w, e.stream, d->vtx_waves, d->panel_rows, nullptr, nullptr);
  dfn::launch_finalize(d->fin_first.as<int>(), d->fin_idx.as<int>(), d->perm.as<int>(), d->ypart.as<double>(), N,
                  d->nops, 1.0, dT, d->nops > 1 ? dN : dT, e.stream);
  DFN_HIP(hipMemcpyAsync(yT.data(), op == 0 ? dT : dN, N * sizeof(double), hipMemcpyDeviceToHost, e.stream));
  DFN_HIP(hipStreamSynchronize(e.stream));
  for (int i = 0; i < N; i++) {
    const double hx = yT[perm[i]];
    y[i] = alpha * hx + (beta == 0.0 ? 0.0 : beta * y[i]);
  }
  return DFN_OK;
  DFN_API_END
}

API int dfn_friction_law(dfn_handle h, double x, const double *Yh, double *Ydh) {
  DFN_API_BEGIN
  if (!h || !Yh || !Ydh) return DFN_ERR_ARG;
  Engine &e = h->e;
  if (e.form != DFN_FORM_REFERENCE) return DFN_ERR_STATE;
  const size_t n3 = 3 * (size_t)e.hp.N;
  DFN_HIP(hipMemcpyAsync(e.scratch3.p, Yh, n3 * sizeof(double), hipMemcpyHostToDevice, e.stream));
  DFN_TRY(e.friction_law_dev(x, e.scratch3.as<double>(), e.scratch3b.as<double>()));
  DFN_HIP(hipMemcpy(Ydh, e.scratch3b.p, n3 * sizeof(double), hipMemcpyDeviceToHost));
  return DFN_OK;
  DFN_API_END
}

API int dfn_compute_P(dfn_handle h, double dt, double *Pth, double *Pdh, int32_t *ier) {
  DFN_API_BEGIN
  if (!h) return DFN_ERR_ARG;
  Engine &e = h->e;
  const size_t N = e.hp.N;
  DFN_TRY(e.cg_launch(&dt, 1, 0));
  int fail = 0;
  DFN_TRY(e.cg_collect(0, 1, &fail));
  if (ier) *ier = fail;
  if (!fail) {
    if (Pth) DFN_HIP(hipMemcpy(Pth, e.Pt(0), N * sizeof(double), hipMemcpyDeviceToHost));
    if (Pdh) DFN_HIP(hipMemcpy(Pdh, e.Pd(0), N * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    e.last_pd_slot = -1;  // the reference leaves P_dot untouched on failure
  }
  return DFN_OK;
  DFN_API_END
}

API int dfn_rkqs(dfn_handle h, double *y, const double *dydx, double *x, double htry, double eps,
                 const double *yscal, double *hdid, double *hnext) {
  DFN_API_BEGIN
  if (!h || !y || !dydx || !x || !yscal || !hdid || !hnext) return DFN_ERR_ARG;
  Engine &e = h->e;
  if (e.form != DFN_FORM_REFERENCE) return DFN_ERR_STATE;
  const size_t n3 = 3 * (size_t)e.hp.N;
  // run on a copy of the handle's state slot: Y <- y
  DFN_HIP(hipMemcpyAsync(e.Y.p, y, n3 * sizeof(double), hipMemcpyHostToDevice, e.stream));
  DFN_HIP(hipMemcpyAsync(e.k(0), dydx, n3 * sizeof(double), hipMemcpyHostToDevice, e.stream));
  DFN_HIP(hipMemcpyAsync(e.yscal.p, yscal, n3 * sizeof(double), hipMemcpyHostToDevice, e.stream));
  DFN_TRY(e.rkqs_dev(*x, htry, eps, hdid, hnext));
  *x = *x + *hdid;
  DFN_HIP(hipStreamSynchronize(e.stream));
  DFN_HIP(hipMemcpy(y, e.Y.p, n3 * sizeof(double), hipMemcpyDeviceToHost));
  return DFN_OK;
  DFN_API_END
}

API int dfn_solve(dfn_handle h, int64_t nsteps, double t_final, dfn_run_stats *stats) {
  DFN_API_BEGIN
  if (!h || nsteps < 0) return DFN_ERR_ARG;
  Engine &e = h->e;
  if (e.hp.kernel == DFN_KERNEL_HMATRIX && !e.active_ops()) return DFN_ERR_STATE;
  int st = DFN_OK;
  for (int64_t i = 0; i < nsteps; i++) {
    if (t_final > 0) {  // land on t_final: the trial step never passes it
      if (e.time >= t_final) break;
      if (e.time + e.t_try > t_final) e.t_try = t_final - e.time;
    }
    st = e.step();
    if (st != DFN_OK) break;
  }
  if (stats) {
    int s2 = dfn_get_stats(h, stats);
    if (st == DFN_OK) st = s2;
  }
  return st;
  DFN_API_END
}

API int dfn_set_output(dfn_handle h, const char *dir, int64_t stride_time, int64_t freq_writing_file) {
  DFN_API_BEGIN
  if (!h) return DFN_ERR_ARG;
  if (dir && h->e.form != DFN_FORM_REFERENCE) return DFN_ERR_STATE;
  return h->e.output_set(dir, stride_time, freq_writing_file);
  DFN_API_END
}

API int dfn_flush_output(dfn_handle h) {
  DFN_API_BEGIN
  if (!h) return DFN_ERR_ARG;
  return h->e.output_write();
  DFN_API_END
}

API int dfn_monitor(dfn_handle h, int32_t on) {
  DFN_API_BEGIN
  if (!h) return DFN_ERR_ARG;
  if (!on) { h->e.monitor = false; return DFN_OK; }
  if (h->e.form != DFN_FORM_REFERENCE) return DFN_ERR_STATE;
  return h->e.monitor_start();
  DFN_API_END
}

API int dfn_monitor_take(dfn_handle h, int64_t cap, double *times, double *maxV, int64_t *count) {
  DFN_API_BEGIN
  if (!h || !count || cap < 0) return DFN_ERR_ARG;
  Engine &e = h->e;
  const int nf = e.hp.nfault;
  const int64_t avail = (int64_t)e.mon_t.size();
  if (!times || !maxV) { *count = avail; return DFN_OK; }
  const int64_t n = std::min(cap, avail);
  std::copy(e.mon_t.begin(), e.mon_t.begin() + n, times);
  std::copy(e.mon_v.begin(), e.mon_v.begin() + n * nf, maxV);
  e.mon_t.erase(e.mon_t.begin(), e.mon_t.begin() + n);
  e.mon_v.erase(e.mon_v.begin(), e.mon_v.begin() + n * nf);
  *count = n;
  return DFN_OK;
  DFN_API_END
}

API int dfn_moment(dfn_handle h, double *M0, double *slip) {
  DFN_API_BEGIN
  if (!h) return DFN_ERR_ARG;
  Engine &e = h->e;
  if (!e.monitor && e.slip.p == nullptr) return DFN_ERR_STATE;
  const int N = e.hp.N;
  std::vector<double> s(N);
  DFN_HIP(hipStreamSynchronize(e.stream));
  DFN_HIP(hipMemcpy(s.data(), e.slip.p, sizeof(double) * N, hipMemcpyDeviceToHost));
  if (slip) std::copy(s.begin(), s.end(), slip);
  if (M0)
    for (int f = 0; f < e.hp.nfault; f++) {  // wrapper/network.m:474-484: sum(mu * slip .* ds)
      double m = 0.0;
      for (int64_t q = 0; q < e.hp.fnel[f]; q++) {
        const int64_t k = e.hp.map1[f] - 1 + q;
        m = m + e.hp.mu * s[k] * e.hp.ds[k];
      }
      M0[f] = m;
    }
  return DFN_OK;
  DFN_API_END
}

API int dfn_kernel_timing(dfn_handle h, int32_t on) {
  DFN_API_BEGIN
  if (!h) return DFN_ERR_ARG;
  Engine &e = h->e;
  DFN_TRY(e.kt_collect());
  e.kt_on = on != 0;
  e.kt_ms = 0.0;
  e.kt_count = 0;
  return DFN_OK;
  DFN_API_END
}

API int dfn_kernel_timing_read(dfn_handle h, double *ms_mean, int64_t *count) {
  DFN_API_BEGIN
  if (!h || !ms_mean || !count) return DFN_ERR_ARG;
  Engine &e = h->e;
  DFN_TRY(e.kt_collect());
  *count = e.kt_count;
  *ms_mean = e.kt_count ? e.kt_ms / (double)e.kt_count : 0.0;
  return DFN_OK;
  DFN_API_END
}

API int dfn_bench_kernel(dfn_handle h, int32_t reps, double *ms) {
  DFN_API_BEGIN
  if (!h || !ms || reps < 1) return DFN_ERR_ARG;
  Engine &e = h->e;
  const double *dV = e.Y.as<double>();
  DFN_TRY(e.apply_kernel(dV, e.sT.as<double>(), e.sN.as<double>()));  // warm
  DFN_HIP(hipStreamSynchronize(e.stream));
  DFN_HIP(hipEventRecord(e.ev0, e.stream));
  for (int r = 0; r < reps; r++) DFN_TRY(e.apply_kernel(dV, e.sT.as<double>(), e.sN.as<double>()));
  DFN_HIP(hipEventRecord(e.ev1, e.stream));
  DFN_HIP(hipEventSynchronize(e.ev1));
  float t = 0;
  DFN_HIP(hipEventElapsedTime(&t, e.ev0, e.ev1));
  ms[0] = (double)t / reps;
  ms[1] = 0.0;
  ms[2] = 0.0;
  ms[3] = 0.0;
  const dfn::DevOps *d = e.active_ops();
  if (e.hp.kernel == DFN_KERNEL_HMATRIX && d) {
    // the dominant launch alone: the phase-2 panels, or phase 1 with the near-field
    // panels when they run as one launch
    double *w = d->w.as<double>();
    DFN_HIP(hipEventRecord(e.ev0, e.stream));
    for (int r = 0; r < reps; r++) {
      if (d->fused)
        dfn::launch_vtx_near(d->vtasks.as<dfn::VTask>(), d->nvt, static_cast<const float *>(d->vdata.p),
                             d->ptasks.as<dfn::PTask>(), d->npt, static_cast<const double *>(d->pdata.p),
                             d->widx.as<int>(), w, d->zpart.as<double>(), d->ypart.as<double>(), e.stream,
                             d->panel_rows, nullptr, d->kc);
      else
        dfn::launch_panels(*d, w, e.stream, d->panel_rows);
    }
    DFN_HIP(hipEventRecord(e.ev1, e.stream));
    DFN_HIP(hipEventSynchronize(e.ev1));
    DFN_HIP(hipEventElapsedTime(&t, e.ev0, e.ev1));
    ms[1] = (double)t / reps;
    ms[2] = d->alg_bytes;
    ms[3] = d->dom_bytes();
  } else {
    // the direct kernels are FP64-arithmetic bound: report algorithmic flops per
    // application, counting sqrt and divide as one each (Mode III: 11 per
    // element-node pair; Mode II: 40 per pair for the two outputs)
    ms[1] = ms[0];
    const double N = e.hp.N;
    ms[2] = (e.hp.mode == DFN_MODE_III ? 11.0 : 40.0) * N * (N + 1.0);
    ms[3] = ms[2];
  }
  return DFN_OK;
  DFN_API_END
}
