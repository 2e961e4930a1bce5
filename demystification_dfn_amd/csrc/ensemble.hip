// Ensemble engine: many independent fault networks advanced together on one GPU.
//
// The networks are concatenated element-wise (network b owns elements [base_b,
// base_b + N_b)); their stress-transfer operators become one block-diagonal
// operator whose two-phase task lists cover every network, so a stage of the
// Runge-Kutta step is one H-matvec launch sequence and one stage kernel for the
// whole ensemble.  Each network keeps its own clock, trial step, step size and
// error norm: the stage kernels read the per-network step size, reduce the
// scaled error per network, and skip networks that have already accepted their
// step (masked rejection: only networks whose attempt was rejected redo stages
// 2..6, with their own reduced step).  compute_P runs as one launch over every
// (network, stage) system.  The per-attempt launch sequences read all their
// parameters from device memory and are captured once as HIP graphs and
// replayed.  Every network follows the operation sequence of the single-network
// engine (engine.hip), so each one's trajectory is the one it would have alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "dfn_device.hpp"
#include "dfn_host.hpp"
#include "engine_common.hpp"

namespace dfn {

namespace {

constexpr int kEnsSlots = 6;

// Concatenates per-network operators into one block-diagonal operator over the
// concatenated element order (H positions and elements of network b offset by
// base[b]).
void concat_ops(const std::vector<HostOperator> &ops, const std::vector<int> &base, HostOperator &out) {
  out = HostOperator();
  int64_t nd = 0, dd = 0, nl = 0, lu = 0, lv = 0;
  for (const HostOperator &H : ops) {
    out.n += H.n;
    out.nbclus += H.nbclus;
    out.nbblclus += H.nbblclus;
    nd += (int64_t)H.d_m.size();
    dd += (int64_t)H.d_data.size();
    nl += (int64_t)H.l_m.size();
    lu += (int64_t)H.l_u.size();
    lv += (int64_t)H.l_v.size();
  }
  out.perm.reserve(out.n);
  out.d_r0.reserve(nd); out.d_c0.reserve(nd); out.d_m.reserve(nd); out.d_n.reserve(nd); out.d_off.reserve(nd);
  out.d_data.reserve(dd);
  out.l_r0.reserve(nl); out.l_c0.reserve(nl); out.l_m.reserve(nl); out.l_n.reserve(nl); out.l_k.reserve(nl);
  out.l_uoff.reserve(nl); out.l_voff.reserve(nl);
  out.l_u.reserve(lu);
  out.l_v.reserve(lv);
  for (size_t b = 0; b < ops.size(); b++) {
    const HostOperator &H = ops[b];
    const int o = base[b];
    for (int p : H.perm) out.perm.push_back(p + o);
    const int64_t doff = (int64_t)out.d_data.size(), uoff = (int64_t)out.l_u.size(), voff = (int64_t)out.l_v.size();
    for (size_t i = 0; i < H.d_m.size(); i++) {
      out.d_r0.push_back(H.d_r0[i] + o); out.d_c0.push_back(H.d_c0[i] + o);
      out.d_m.push_back(H.d_m[i]); out.d_n.push_back(H.d_n[i]); out.d_off.push_back(H.d_off[i] + doff);
    }
    out.d_data.insert(out.d_data.end(), H.d_data.begin(), H.d_data.end());
    for (size_t i = 0; i < H.l_m.size(); i++) {
      out.l_r0.push_back(H.l_r0[i] + o); out.l_c0.push_back(H.l_c0[i] + o);
      out.l_m.push_back(H.l_m[i]); out.l_n.push_back(H.l_n[i]); out.l_k.push_back(H.l_k[i]);
      out.l_uoff.push_back(H.l_uoff[i] + uoff); out.l_voff.push_back(H.l_voff[i] + voff);
    }
    out.l_u.insert(out.l_u.end(), H.l_u.begin(), H.l_u.end());
    out.l_v.insert(out.l_v.end(), H.l_v.begin(), H.l_v.end());
    for (size_t i = 0; i < H.leaf_r0.size(); i++) {
      out.leaf_r0.push_back(H.leaf_r0[i] + o);
      out.leaf_m.push_back(H.leaf_m[i]);
    }
  }
}

template <class T>
void append(std::vector<T> &dst, const std::vector<T> &src) {
  dst.insert(dst.end(), src.begin(), src.end());
}

}  // namespace

struct EnsEngine {
  int B = 0, Ntot = 0, max_n = 0, mode = DFN_MODE_II, nops = 2;
  double kscale = 0.0;
  std::vector<HostProblem> hp;
  std::vector<int> base, xbase;
  std::vector<HostOperator> ops[2];  // per network until the layout is built
  bool built = false;
  int precision = DFN_PREC_FP64;
  bool use_graph = true;
  hipStream_t stream = nullptr;
  hipStream_t cg_stream = nullptr;  // the first attempt's pressure solves run beside the k1 stress evaluation
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  std::unique_ptr<DevOps> ops64, ops32;  // FP64 layout; the reduced-precision layout of kind ops32_kind
  int ops32_kind = -1;
  HostOperator cat[2];  // concatenated operators (kept for the FP32 layout)

  // concatenated problem data
  DMem a, b, Dc, f0, V0, nld, sld, fake, net, ds, perm, permx, src_elem, src_Q, src_tb, src_te, cross, xrow_start,
      xrow_ent, cgnets;
  // state and stage buffers
  DMem Y, Ynew, yscal, sT, sN, kbuf, ytA, ytB, P, Pdot;
  // per-network parameters read by the captured launches
  DMem d_h, d_ttry, d_tol, d_eta, d_active, status, cg_work, cg_res;
  // compute_P systems, (network, slot) with the slots of a network consecutive:
  // networks of at most kCgLdsMax elements (one workgroup per system, vectors in
  // registers) and larger ones (the cooperative kernel, each system over
  // ceil(N / kCgSpan) workgroups, or the plain one-workgroup kernel), each list for
  // all six slots and for the five stage slots of a retried attempt
  DMem sys_small_all, sys_small_stage, sys_large_all, sys_large_stage;
  int n_small = 0, n_large = 0, max_small = 0, max_large = 0;
  // cooperative launches of the large networks' systems: groups of whole systems
  // whose workgroups can all be resident; per group, its map entries
  bool coop = false;
  int coop_cap = 0;  // most crossing entries in one span of any large network
  struct CoopGroup { int first, nwg; };
  std::vector<CoopGroup> coop_all, coop_stage;
  DMem map_all, map_stage, coop_records;
  unsigned cg_seq = 0;
  // Launches issued one by one with cooperative pressure solves: the groups hold the
  // systems slot by slot (every network's slot 0 first), an event follows each
  // group and the small networks' launch, and the stage kernel that reads slot s
  // waits only for the group that completes slot s, so the later solves run beside
  // the stress evaluations of the earlier stages.
  std::vector<int> last_group_all, last_group_stage;  // per list slot: the group completing it
  std::vector<hipEvent_t> ev_group;
  hipEvent_t ev_small = nullptr;
  const std::vector<int> *cur_last = nullptr;  // the last cg_launch's list
  int cur_first = 0;                           // its first slot
  bool per_slot() const { return coop && !use_graph; }
  // pinned host mirrors
  double *h_h = nullptr, *h_ttry = nullptr;
  int *h_active = nullptr, *h_cgres = nullptr;
  StatusWord *h_status = nullptr;
  CgNet *h_nets = nullptr;
  std::vector<CgNet> nets;
  // per-network clocks and counters
  std::vector<double> time, t_try;
  std::vector<dfn_run_stats> st;
  hipGraphExec_t g_try = nullptr, g_retry = nullptr;
  const DevOps *graph_ops = nullptr;
  // live timing of the phase-2 panel kernel (dfn_ensemble_kernel_timing): an event
  // pair around each panel launch of an evaluation in which every network takes
  // part (launches issued one by one while it is on)
  bool kt_on = false, kt_graph = true, all_active = false;
  std::vector<hipEvent_t> kt_ev;
  int kt_used = 0;
  double kt_ms = 0.0;
  int64_t kt_count = 0;

  // step-size control: 0 = rkqs's proposal (the reference), 1 = PI controller,
  // 2 = the proposal network with the PI controller as fallback
  int controller = 0;
  Mlp *mlp = nullptr;
  bool recording = false;
  std::vector<float> rec_X, rec_T;  // (features, log10 accepted-step ratio) pairs
  std::vector<double> prev_err, prev_err2, prev_lmax, pend_href;
  std::vector<int> fallback, learned_last, has_pend, pend_skip;
  std::vector<float> pend_feat;
  DMem vstats, feat_d, out_d;
  float *h_feat = nullptr, *h_out = nullptr;
  double *h_vstats = nullptr;  // per network: max |V|, mean log10 |V|, fraction above 1e-3 m/s
  static constexpr int kFallbackSteps = 3;

  double *kk(int i) const { return kbuf.as<double>() + (size_t)i * 3 * Ntot; }
  double *cg_vec(int slot, int which) const { return cg_work.as<double>() + ((size_t)slot * 12 + which) * Ntot; }

  ~EnsEngine() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (hipEvent_t ev : kt_ev) (void)hipEventDestroy(ev);
    if (cg_stream) (void)hipStreamSynchronize(cg_stream);
    for (hipEvent_t ev : ev_group) (void)hipEventDestroy(ev);
    if (ev_small) (void)hipEventDestroy(ev_small);
    drop_graphs();
    if (h_h) (void)hipHostFree(h_h);
    if (h_ttry) (void)hipHostFree(h_ttry);
    if (h_active) (void)hipHostFree(h_active);
    if (h_cgres) (void)hipHostFree(h_cgres);
    if (h_status) (void)hipHostFree(h_status);
    if (h_nets) (void)hipHostFree(h_nets);
    if (h_feat) (void)hipHostFree(h_feat);
    if (h_out) (void)hipHostFree(h_out);
    if (h_vstats) (void)hipHostFree(h_vstats);
    ops64.reset();
    ops32.reset();
    if (cg_stream) (void)hipStreamSynchronize(cg_stream);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (cg_stream) (void)hipStreamDestroy(cg_stream);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void drop_graphs() {
    if (g_try) (void)hipGraphExecDestroy(g_try);
    if (g_retry) (void)hipGraphExecDestroy(g_retry);
    g_try = g_retry = nullptr;
    graph_ops = nullptr;
  }

  int init(const dfn_problem *probs, int nnet) {
    B = nnet;
    hp.resize(B);
    for (int i = 0; i < B; i++) DFN_TRY(setup_problem(&probs[i], hp[i]));
    mode = hp[0].mode;
    nops = mode == DFN_MODE_II ? 2 : 1;
    const double pi = 4.0 * std::atan(1.0);
    auto scale_of = [&](const HostProblem &p) {
      return p.mode == DFN_MODE_III ? p.mu / (2.0 * pi) : p.mu / pi * (1.0 - p.cs * p.cs / (p.cp * p.cp));
    };
    kscale = scale_of(hp[0]);
    for (const HostProblem &p : hp) {
      // one block-diagonal operator with one output scale; the step batches the
      // compute_P solves of a step, which needs a permeability that does not evolve
      if (p.mode != mode || p.kernel != DFN_KERNEL_HMATRIX || scale_of(p) != kscale || p.coupling == 1)
        return DFN_ERR_UNSUPPORTED;
    }
    base.resize(B);
    xbase.resize(B);
    std::vector<double> ca, cb, cDc, cf0, cV0, cnld, csld, cds, cperm, cpx, cQ, ctb, cte, cP;
    std::vector<uint8_t> cfake;
    std::vector<int> cnet;
    std::vector<int64_t> csrc, ccross;
    std::vector<int> xstart, xent;
    nets.resize(B);
    int off = 0, xo = 0, so = 0, co = 0;
    for (int i = 0; i < B; i++) {
      const HostProblem &p = hp[i];
      const int N = p.N;
      base[i] = off;
      xbase[i] = xo;
      append(ca, p.a); append(cb, p.b); append(cDc, p.Dc); append(cf0, p.f0); append(cV0, p.V0);
      append(cnld, p.nld); append(csld, p.sld); append(cfake, p.is_fake); append(cds, p.ds);
      append(cperm, p.perm0); append(cP, p.P0);
      for (int k = 0; k < N; k++) cnet.push_back(i);
      // permeability_x (initialisation.f90:252-266), N + 1 entries per network
      std::vector<double> px(N + 1, 0.0);
      for (int k = 1; k < N; k++) px[k] = 2.0 * p.perm0[k - 1] * p.perm0[k] / (p.perm0[k - 1] + p.perm0[k]);
      append(cpx, px);
      append(csrc, p.src_elem); append(cQ, p.src_Q); append(ctb, p.src_tb); append(cte, p.src_te);
      append(ccross, p.cross);
      // crossing rows: entry indices global, crossing ids network-local
      std::vector<std::vector<std::pair<int, int>>> rows(N);
      for (int c = 0; c < p.ncross(); c++) {
        const int64_t n1 = p.cross[6 * c + 2], n2 = p.cross[6 * c + 3];
        const int64_t e[4] = {n1 - 2, n1 - 1, n2 - 2, n2 - 1};
        for (int z = 0; z < 4; z++) rows[e[z]].push_back({c, z});
      }
      for (int k = 0; k < N; k++) {
        xstart.push_back((int)(xent.size() / 2));
        for (auto [c, z] : rows[k]) { xent.push_back(c); xent.push_back(z); }
      }
      CgNet &c = nets[i];
      c.n = N; c.base = off; c.xbase = xo; c.src0 = so; c.nsrc = (int)p.src_elem.size();
      c.cross0 = co; c.ncross = p.ncross(); c.pad = 0;
      c.cfac = p.por * p.visc * (p.fcomp + p.rcomp);
      c.time = 0.0; c.tol = p.tol; c.por = p.por; c.beta = p.fcomp + p.rcomp;
      off += N;
      xo += N + 1;
      so += c.nsrc;
      co += c.ncross;
      max_n = std::max(max_n, N);
    }
    Ntot = off;
    xstart.push_back((int)(xent.size() / 2));
    if (xent.empty()) xent.assign(2, 0);
    // state in the ensemble layout [V | psi | sigmaN], each over all networks
    std::vector<double> y3((size_t)3 * Ntot);
    for (int i = 0; i < B; i++)
      for (int c = 0; c < 3; c++)
        std::copy(hp[i].Y0.begin() + (size_t)c * hp[i].N, hp[i].Y0.begin() + (size_t)(c + 1) * hp[i].N,
                  y3.begin() + (size_t)c * Ntot + base[i]);
    DFN_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    DFN_HIP(hipStreamCreateWithFlags(&cg_stream, hipStreamNonBlocking));
    DFN_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    DFN_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    DFN_TRY(a.upload(ca)); DFN_TRY(b.upload(cb)); DFN_TRY(Dc.upload(cDc)); DFN_TRY(f0.upload(cf0));
    DFN_TRY(V0.upload(cV0)); DFN_TRY(nld.upload(cnld)); DFN_TRY(sld.upload(csld)); DFN_TRY(fake.upload(cfake));
    DFN_TRY(net.upload(cnet)); DFN_TRY(ds.upload(cds)); DFN_TRY(perm.upload(cperm)); DFN_TRY(permx.upload(cpx));
    DFN_TRY(src_elem.upload(csrc)); DFN_TRY(src_Q.upload(cQ)); DFN_TRY(src_tb.upload(ctb));
    DFN_TRY(src_te.upload(cte)); DFN_TRY(cross.upload(ccross)); DFN_TRY(xrow_start.upload(xstart));
    DFN_TRY(xrow_ent.upload(xent)); DFN_TRY(cgnets.upload(nets));
    const size_t n3 = (size_t)3 * Ntot * sizeof(double);
    DFN_TRY(Y.upload(y3));
    DFN_TRY(Ynew.alloc(n3)); DFN_TRY(yscal.alloc(n3)); DFN_TRY(kbuf.alloc(6 * n3));
    DFN_TRY(ytA.alloc(n3)); DFN_TRY(ytB.alloc(n3));
    DFN_TRY(sT.alloc(sizeof(double) * Ntot)); DFN_TRY(sN.alloc(sizeof(double) * Ntot));
    DFN_HIP(hipMemset(sN.p, 0, sizeof(double) * Ntot));
    DFN_TRY(P.upload(cP));
    DFN_TRY(Pdot.alloc(sizeof(double) * Ntot));
    DFN_HIP(hipMemset(Pdot.p, 0, sizeof(double) * Ntot));
    DFN_TRY(cg_work.alloc(sizeof(double) * (size_t)kEnsSlots * 12 * Ntot));
    DFN_TRY(cg_res.alloc(sizeof(int) * 2 * kEnsSlots * B));
    DFN_TRY(d_h.alloc(sizeof(double) * B)); DFN_TRY(d_ttry.alloc(sizeof(double) * B));
    DFN_TRY(d_active.alloc(sizeof(int) * B)); DFN_TRY(status.alloc(sizeof(StatusWord) * B));
    std::vector<double> tol(B), eta(B);
    for (int i = 0; i < B; i++) { tol[i] = hp[i].tol; eta[i] = hp[i].mu / (2.0 * hp[i].cs); }
    DFN_TRY(d_tol.upload(tol));
    DFN_TRY(d_eta.upload(eta));
    DFN_TRY(setup_cg_systems(xstart));
    DFN_HIP(hipHostMalloc((void **)&h_h, sizeof(double) * B, hipHostMallocDefault));
    DFN_HIP(hipHostMalloc((void **)&h_ttry, sizeof(double) * B, hipHostMallocDefault));
    DFN_HIP(hipHostMalloc((void **)&h_active, sizeof(int) * B, hipHostMallocDefault));
    DFN_HIP(hipHostMalloc((void **)&h_cgres, sizeof(int) * 2 * kEnsSlots * B, hipHostMallocDefault));
    DFN_HIP(hipHostMalloc((void **)&h_status, sizeof(StatusWord) * B, hipHostMallocDefault));
    DFN_HIP(hipHostMalloc((void **)&h_nets, sizeof(CgNet) * B, hipHostMallocDefault));
    time.assign(B, 0.0);
    t_try.assign(B, 100.0);
    st.assign(B, dfn_run_stats{});
    DFN_TRY(vstats.alloc(sizeof(double) * 3 * B));
    DFN_TRY(feat_d.alloc(sizeof(float) * DFN_MLP_INPUTS * B));
    DFN_TRY(out_d.alloc(sizeof(float) * B));
    DFN_HIP(hipHostMalloc((void **)&h_feat, sizeof(float) * DFN_MLP_INPUTS * B, hipHostMallocDefault));
    DFN_HIP(hipHostMalloc((void **)&h_out, sizeof(float) * B, hipHostMallocDefault));
    DFN_HIP(hipHostMalloc((void **)&h_vstats, sizeof(double) * 3 * B, hipHostMallocDefault));
    prev_err.assign(B, -1.0);   // < 0: no earlier accepted step
    prev_err2.assign(B, -1.0);
    prev_lmax.assign(B, 0.0);
    pend_href.assign(B, 0.0);
    fallback.assign(B, 0);
    learned_last.assign(B, 0);
    has_pend.assign(B, 0);
    pend_skip.assign(B, 0);
    pend_feat.assign((size_t)DFN_MLP_INPUTS * B, 0.0f);
    return DFN_OK;
  }

  // compute_P system lists and the cooperative launch groups (see the members)
  int setup_cg_systems(const std::vector<int> &xstart) {
    DFN_HIP(hipMemset(cg_work.p, 0, sizeof(double) * (size_t)kEnsSlots * 12 * Ntot));  // message epochs 0
    // DFN_CG_PATH=plain (read at creation, as the single-network engine does): the
    // plain one-workgroup kernel for every network
    const char *cp = std::getenv("DFN_CG_PATH");
    const bool plain = cp && std::string(cp) == "plain";
    std::vector<CgSys> sa, ss, la, ls;
    std::vector<int> large;
    for (int i = 0; i < B; i++) {
      const bool small = !plain && hp[i].N <= kCgLdsMax;
      if (small) { n_small++; max_small = std::max(max_small, hp[i].N); }
      else { n_large++; max_large = std::max(max_large, hp[i].N); large.push_back(i); }
      for (int sl = 0; sl < kEnsSlots; sl++) {
        CgSys c{};
        c.dt = 0.0;
        c.diag = cg_vec(sl, 0) + base[i]; c.A1 = cg_vec(sl, 1) + base[i]; c.A2 = cg_vec(sl, 2) + base[i];
        c.A3 = cg_vec(sl, 3) + base[i]; c.b = cg_vec(sl, 4) + base[i]; c.x = cg_vec(sl, 5) + base[i];
        c.r = cg_vec(sl, 6) + base[i]; c.p = cg_vec(sl, 7) + base[i]; c.Ap = cg_vec(sl, 8) + base[i];
        c.Pt = cg_vec(sl, 10) + base[i]; c.Pd = cg_vec(sl, 11) + base[i];
        c.result = cg_res.as<int>() + 2 * ((size_t)i * kEnsSlots + sl);
        c.net = i;
        (small ? sa : la).push_back(c);
        if (sl > 0) (small ? ss : ls).push_back(c);
      }
    }
    DFN_TRY(sys_small_all.upload(sa));
    DFN_TRY(sys_small_stage.upload(ss));
    DFN_TRY(sys_large_all.upload(la));
    DFN_TRY(sys_large_stage.upload(ls));
    if (n_large == 0) return DFN_OK;
    // the cooperative kernel for the large networks when every span's crossing
    // entries fit its LDS and one network's workgroups can be resident together
    coop_cap = 0;
    for (int i : large)
      for (int s0 = 0; s0 < hp[i].N; s0 += kCgSpan)
        coop_cap = std::max(coop_cap, xstart[base[i] + std::min(s0 + kCgSpan, hp[i].N)] - xstart[base[i] + s0]);
    const int capacity = plain ? 0 : cg_coop_capacity(coop_cap);
    int max_wps = 0;
    for (int i : large) max_wps = std::max(max_wps, (hp[i].N + kCgSpan - 1) / kCgSpan);
    coop = capacity > 0 && max_wps <= capacity;
    if (!coop) return DFN_OK;
    // whole systems packed slot by slot (every large network's first slot, then its
    // second, ...) into launches of at most `capacity` workgroups; the list itself
    // keeps each network's slots together (system index li * nslot + sl)
    auto plan = [&](int nslot, std::vector<CoopGroup> &groups, DMem &map, std::vector<int> &last) -> int {
      std::vector<int4> m;
      groups.clear();
      last.assign(nslot, 0);
      int in_group = 0, group_first = 0;
      for (int sl = 0; sl < nslot; sl++) {
        for (size_t li = 0; li < large.size(); li++) {
          const int i = large[li], sysi = (int)li * nslot + sl;
          const int wps = (hp[i].N + kCgSpan - 1) / kCgSpan;
          if (in_group + wps > capacity) {
            groups.push_back({group_first, in_group});
            group_first = (int)m.size();
            in_group = 0;
          }
          for (int q = 0; q < wps; q++) m.push_back(make_int4(sysi, q, wps, in_group));
          in_group += wps;
        }
        last[sl] = (int)groups.size();  // the group being filled holds this slot's last system
      }
      groups.push_back({group_first, in_group});
      return map.upload(m);
    };
    DFN_TRY(plan(kEnsSlots, coop_all, map_all, last_group_all));
    DFN_TRY(plan(kEnsSlots - 1, coop_stage, map_stage, last_group_stage));
    ev_group.assign(std::max(coop_all.size(), coop_stage.size()), nullptr);
    for (hipEvent_t &ev : ev_group) DFN_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    DFN_HIP(hipEventCreateWithFlags(&ev_small, hipEventDisableTiming));
    // message records: kCgRecordWords per workgroup of the largest launch
    int most = 0;
    for (const CoopGroup &g : coop_all) most = std::max(most, g.nwg);
    DFN_TRY(coop_records.alloc(sizeof(unsigned long long) * kCgRecordWords * (size_t)most));
    DFN_HIP(hipMemset(coop_records.p, 0, sizeof(unsigned long long) * kCgRecordWords * (size_t)most));
    use_graph = false;  // cooperative launches are issued one by one, not captured
    return DFN_OK;
  }

  // PI step-size controller (Gustafsson): exponents 0.7/5 on the current and 0.4/5
  // on the previous scaled error, growth within [0.2, 5] per step
  static double pi_next(double h, double err, double err_prev) {
    const double e = std::fmax(err, 1e-10), ep = std::fmax(err_prev, 1e-10);
    double fac = 0.9 * std::pow(e, -0.7 / 5.0) * std::pow(ep, 0.4 / 5.0);
    fac = std::fmin(5.0, std::fmax(0.2, fac));
    return h * fac;
  }

  // the proposal network's features of network i after an accepted step of size
  // hdid with scaled error err: max and mean log10 |V|, the rate of change of max
  // log10 |V| over the step, the scaled errors of the last three accepted steps
  // (log10; the oldest repeat the newest at the start), log10 of the step and the
  // fraction of elements slipping faster than 1e-3 m/s
  void features(int i, double hdid, double err, const double *vs, float *f) const {
    const double lmax = std::log10(std::fmax(vs[0], 1e-30));
    const bool first = prev_err[i] < 0.0;
    const double e0 = std::log10(std::fmax(err, 1e-10));
    const double e1 = first ? e0 : std::log10(std::fmax(prev_err[i], 1e-10));
    const double e2 = prev_err2[i] < 0.0 ? e1 : std::log10(std::fmax(prev_err2[i], 1e-10));
    f[0] = (float)lmax;
    f[1] = (float)vs[1];
    f[2] = first ? 0.0f : (float)((lmax - prev_lmax[i]) / std::fmax(hdid, 1e-30));
    f[3] = (float)e0;
    f[4] = (float)e1;
    f[5] = (float)e2;
    f[6] = (float)std::log10(std::fmax(hdid, 1e-30));
    f[7] = (float)vs[2];
  }

  // next trial steps from this step's outcome (t_try before the one-day cap)
  int control(const std::vector<double> &hdid, const std::vector<double> &href, const std::vector<double> &err,
              const std::vector<int> &nrej, const std::vector<int> &guarded, const std::vector<int> &stepping,
              const std::vector<int> &clipped) {
    if (controller == 0 && !recording) {
      for (int i = 0; i < B; i++)
        if (stepping[i]) t_try[i] = href[i];
      return DFN_OK;
    }
    launch_ens_vstats(Y.as<double>(), fake.as<uint8_t>(), cgnets.as<CgNet>(), B, vstats.as<double>(), stream);
    DFN_HIP(hipMemcpyAsync(h_vstats, vstats.p, sizeof(double) * 3 * B, hipMemcpyDeviceToHost, stream));
    DFN_HIP(hipStreamSynchronize(stream));
    std::vector<int> use_mlp(B, 0);
    for (int i = 0; i < B; i++) {
      if (!stepping[i]) continue;
      float *f = h_feat + (size_t)i * DFN_MLP_INPUTS;
      features(i, hdid[i], err[i], h_vstats + 3 * (size_t)i, f);
      if (recording) {
        if (has_pend[i] && !pend_skip[i] && !guarded[i] && !clipped[i]) {
          rec_X.insert(rec_X.end(), pend_feat.begin() + (size_t)i * DFN_MLP_INPUTS,
                       pend_feat.begin() + (size_t)(i + 1) * DFN_MLP_INPUTS);
          rec_T.push_back((float)std::log10(hdid[i] / pend_href[i]));
        }
        std::copy(f, f + DFN_MLP_INPUTS, pend_feat.begin() + (size_t)i * DFN_MLP_INPUTS);
        pend_href[i] = href[i];
        // the cap or the landing on a final time, not the controller, sets the next step
        pend_skip[i] = href[i] > 24 * 60 * 60 || clipped[i];
        has_pend[i] = 1;
      }
      if (controller == 2 && learned_last[i] && nrej[i] > 0) fallback[i] = kFallbackSteps;
      if (controller == 0) t_try[i] = href[i];
      else if (controller == 1 || fallback[i] > 0) t_try[i] = pi_next(hdid[i], err[i], prev_err[i]);
      else use_mlp[i] = 1;
      learned_last[i] = use_mlp[i];
      if (controller == 2 && !use_mlp[i] && fallback[i] > 0) fallback[i]--;
      prev_err2[i] = prev_err[i];
      prev_err[i] = err[i];
      prev_lmax[i] = f[0];
    }
    if (controller == 2) {
      if (!mlp) return DFN_ERR_STATE;
      DFN_HIP(hipMemcpyAsync(feat_d.p, h_feat, sizeof(float) * DFN_MLP_INPUTS * B, hipMemcpyHostToDevice, stream));
      DFN_TRY(mlp_forward_device(mlp, feat_d.as<float>(), B, out_d.as<float>(), stream));
      DFN_HIP(hipMemcpyAsync(h_out, out_d.p, sizeof(float) * B, hipMemcpyDeviceToHost, stream));
      DFN_HIP(hipStreamSynchronize(stream));
      for (int i = 0; i < B; i++)
        if (use_mlp[i]) {
          // log10 of the ratio to rkqs's proposal, kept within [1/10, 2]
          const double y = std::fmin(0.30103, std::fmax(-1.0, (double)mlp_output_scale(mlp, h_out[i])));
          t_try[i] = href[i] * std::pow(10.0, y);
        }
    }
    return DFN_OK;
  }

  int build_hmat() {
    const char *hb = std::getenv("DFN_HBUILD");
    const bool on_host = hb && std::string(hb) == "host";
    for (int o = 0; o < nops; o++) ops[o].assign(B, HostOperator());
    for (int i = 0; i < B; i++) {
      // N then T, as the single-network build
      for (int o = nops - 1; o >= 0; o--) {
        const int op = o == 1 ? DFN_OP_N : DFN_OP_T;
        DFN_TRY(on_host ? build_operator(hp[i], op, ops[o][i]) : build_operator_device(hp[i], op, ops[o][i], stream));
      }
    }
    for (int o = 0; o < nops; o++) {
      concat_ops(ops[o], base, cat[o]);
      ops[o].clear();
      ops[o].shrink_to_fit();
    }
    built = true;
    return relayout();
  }

  int relayout() {
    drop_graphs();
    ops64.reset();
    ops32.reset();
    if (!built) return DFN_OK;
    const HostOperator *list[2] = {&cat[0], &cat[1]};
    // the network of each H position: network b owns [base_b, base_b + N_b)
    std::vector<int> hnet(Ntot);
    for (int i = 0; i < B; i++) std::fill(hnet.begin() + base[i], hnet.begin() + base[i] + hp[i].N, i);
    ops64.reset(new DevOps);
    DFN_TRY(layout_ops<double>(list, nops, *ops64, &hnet));
    if (precision != DFN_PREC_FP64) DFN_TRY(layout_reduced());
    return DFN_OK;
  }

  // the reduced-precision layout of the current precision (see Engine::layout_reduced)
  int layout_reduced() {
    drop_graphs();
    ops32.reset();
    ops32_kind = -1;
    if (!built) return DFN_OK;
    const HostOperator *list[2] = {&cat[0], &cat[1]};
    std::vector<int> hnet(Ntot);
    for (int i = 0; i < B; i++) std::fill(hnet.begin() + base[i], hnet.begin() + base[i] + hp[i].N, i);
    ops32.reset(new DevOps);
    const bool split = precision == DFN_PREC_MIXED_SPLIT, mixed = precision == DFN_PREC_MIXED || split;
    DFN_TRY(layout_ops<float>(list, nops, *ops32, &hnet, mixed, precision == DFN_PREC_FP32_STORE || mixed, split));
    ops32_kind = precision;
    return DFN_OK;
  }

  const DevOps *active_ops() const {
    return (precision != DFN_PREC_FP64 && ops32 && ops32_kind == precision) ? ops32.get() : ops64.get();
  }

  hipEvent_t *kt_next() {
    if (!kt_on || !all_active) return nullptr;
    if (kt_used >= 1024 && kt_collect() != DFN_OK) return nullptr;
    while ((int)kt_ev.size() < 2 * (kt_used + 1)) {
      hipEvent_t ev = nullptr;
      if (hipEventCreate(&ev) != hipSuccess) return nullptr;
      kt_ev.push_back(ev);
    }
    return &kt_ev[2 * kt_used++];
  }

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

  void apply(const double *dV) {
    const DevOps *d = active_ops();
    hipEvent_t *kt = kt_next();
    double *w = d->w.as<double>();
    launch_gather(dV, d->perm.as<int>(), w, Ntot, stream, d->xf.as<float>());
    // the tasks of networks that sit out this attempt (accepted, or waiting for
    // others) return at once: a retry streams only the retrying networks' operators
    launch_phases(*d, w, stream, d->vtx_waves, d->panel_rows, d_active.as<int>(), kt);
    launch_finalize(d->fin_first.as<int>(), d->fin_idx.as<int>(), d->perm.as<int>(), d->ypart.as<double>(), Ntot,
                    d->nops, kscale, sT.as<double>(), d->nops > 1 ? sN.as<double>() : sT.as<double>(), stream);
  }

  EnsStageArgs stage_args(const double *ytemp, int slot, double *ynext) {
    EnsStageArgs E{};
    StageArgs &s = E.a;
    s.n = Ntot;
    s.p = RhsParams{a.as<double>(), b.as<double>(), Dc.as<double>(), f0.as<double>(), V0.as<double>(),
                    nld.as<double>(), sld.as<double>(), fake.as<uint8_t>(), 0.0, mode == DFN_MODE_III ? 1 : 0};
    s.y = Y.as<double>();
    s.ytemp = ytemp;
    s.sT = sT.as<double>();
    s.sN = mode == DFN_MODE_II ? sN.as<double>() : nullptr;
    s.Pt = cg_vec(slot, 10);
    s.Pd = cg_vec(slot, 11);
    for (int i = 0; i < 6; i++) s.k[i] = kk(i);
    s.ynext = ynext;
    s.yscal = yscal.as<double>();
    E.net = net.as<int>();
    E.active = d_active.as<int>();
    E.h = d_h.as<double>();
    E.t_try = d_ttry.as<double>();
    E.tol = d_tol.as<double>();
    E.eta = d_eta.as<double>();
    E.status = status.as<StatusWord>();
    return E;
  }

  CgArgs cg_args(const CgSys *sys) {
    CgArgs A{};
    A.n = max_n;
    A.sys = sys;
    A.perm_x = permx.as<double>();
    A.perm_x_diag = permx.as<double>();
    A.perm = perm.as<double>();
    A.perm_diag = perm.as<double>();
    A.ds = ds.as<double>();
    A.P = P.as<double>();
    A.src_elem = src_elem.as<int64_t>();
    A.src_Q = src_Q.as<double>();
    A.src_tb = src_tb.as<double>();
    A.src_te = src_te.as<double>();
    A.cross = cross.as<int64_t>();
    A.nets = cgnets.as<CgNet>();
    A.active = d_active.as<int>();
    return A;
  }

  int cg_launch(bool stages_only, hipStream_t cs) {
    const int nslot = stages_only ? kEnsSlots - 1 : kEnsSlots, first = stages_only ? 1 : 0;
    CrossRows X{xrow_start.as<int>(), xrow_ent.as<int>()};
    if (n_small > 0) {
      CgSys *sys = stages_only ? sys_small_stage.as<CgSys>() : sys_small_all.as<CgSys>();
      launch_ens_cg_dt(sys, n_small, nslot, d_ttry.as<double>(), d_h.as<double>(), first, cs);
      CgArgs A = cg_args(sys);
      A.n = max_small;
      launch_cg_lds_multi(A, X, n_small * nslot, max_small, cs);
      if (per_slot()) DFN_HIP(hipEventRecord(ev_small, cs));
    }
    if (n_large > 0) {
      CgSys *sys = stages_only ? sys_large_stage.as<CgSys>() : sys_large_all.as<CgSys>();
      launch_ens_cg_dt(sys, n_large, nslot, d_ttry.as<double>(), d_h.as<double>(), first, cs);
      CgArgs A = cg_args(sys);
      A.n = max_large;
      if (!coop) {
        launch_cg(A, n_large * nslot, cs);
        return DFN_OK;
      }
      const std::vector<CoopGroup> &groups = stages_only ? coop_stage : coop_all;
      const int4 *map = (stages_only ? map_stage : map_all).as<int4>();
      cur_last = stages_only ? &last_group_stage : &last_group_all;
      cur_first = first;
      for (size_t gi = 0; gi < groups.size(); gi++) {
        const CoopGroup &g = groups[gi];
        if (++cg_seq >= kCgMaxSeq) {  // epochs would repeat: clear every message word first
          DFN_HIP(hipMemsetAsync(coop_records.p, 0, coop_records.bytes, cs));
          DFN_HIP(hipMemsetAsync(cg_work.p, 0, sizeof(double) * (size_t)kEnsSlots * 12 * Ntot, cs));
          cg_seq = 1;
        }
        A.wg_map = map + g.first;
        if (!launch_cg_coop_map(A, X, g.nwg, coop_cap, cg_seq, coop_records.as<unsigned long long>(), cs))
          return DFN_ERR_HIP;
        if (per_slot()) DFN_HIP(hipEventRecord(ev_group[gi], cs));
      }
    }
    return DFN_OK;
  }

  // the stage kernel about to read pressure slot `slot` waits for its solves
  void wait_slot(int slot) {
    if (n_small > 0) (void)hipStreamWaitEvent(stream, ev_small, 0);
    if (n_large > 0 && cur_last) (void)hipStreamWaitEvent(stream, ev_group[(*cur_last)[slot - cur_first]], 0);
  }

  // stages 2..6 of rkck (ode_bs.f90:199-224) for the active networks; with per-slot
  // waits each stage kernel first waits for its own pressure slot
  void attempt(bool waits = false) {
    launch_ens_clear_status(status.as<StatusWord>(), d_active.as<int>(), B, stream);
    double *in = ytA.as<double>(), *out = ytB.as<double>();
    for (int s = 2; s <= 6; s++) {
      apply(in);
      if (waits) wait_slot(s - 1);
      const EnsStageArgs E = stage_args(in, s - 1, s == 6 ? Ynew.as<double>() : out);
      launch_ens_stage(s, E, stream);
      std::swap(in, out);
    }
  }

  // first attempt of a step: compute_P of all six slots, k1 and yscal, stages 2..6
  // (the pressure solves on the side stream, joined before stage 1, which reads them)
  void seq_try() {
    const bool ps = per_slot();
    (void)hipEventRecord(ev_fork, stream);
    (void)hipStreamWaitEvent(cg_stream, ev_fork, 0);
    seq_rc = cg_launch(false, cg_stream);
    if (!ps) (void)hipEventRecord(ev_join, cg_stream);
    apply(Y.as<double>());
    if (ps) wait_slot(0);
    else (void)hipStreamWaitEvent(stream, ev_join, 0);
    launch_ens_stage(1, stage_args(Y.as<double>(), 0, ytA.as<double>()), stream);
    attempt(ps);
  }

  // retried attempt with reduced steps: compute_P of slots 1..5, restart, stages 2..6
  // (with per-slot waits the solves run on the side stream beside the stages)
  void seq_retry() {
    const bool ps = per_slot();
    if (ps) {
      (void)hipEventRecord(ev_fork, stream);
      (void)hipStreamWaitEvent(cg_stream, ev_fork, 0);
    }
    seq_rc = cg_launch(true, ps ? cg_stream : stream);
    launch_ens_restart_stage2(stage_args(Y.as<double>(), 0, ytA.as<double>()), stream);
    attempt(ps);
  }
  int seq_rc = DFN_OK;  // status of the last sequence's compute_P launches

  int capture(void (EnsEngine::*fn)(), hipGraphExec_t *exec) {
    hipGraph_t g = nullptr;
    DFN_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    (this->*fn)();
    DFN_HIP(hipStreamEndCapture(stream, &g));
    const hipError_t e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return DFN_ERR_HIP;
    return DFN_OK;
  }

  int run(bool retry) {
    if (use_graph) {
      if (graph_ops != active_ops()) {
        drop_graphs();
        DFN_TRY(capture(&EnsEngine::seq_try, &g_try));
        DFN_TRY(capture(&EnsEngine::seq_retry, &g_retry));
        graph_ops = active_ops();
      }
      DFN_HIP(hipGraphLaunch(retry ? g_retry : g_try, stream));
    } else {
      if (retry) seq_retry();
      else seq_try();
      DFN_TRY(seq_rc);
      DFN_HIP(hipGetLastError());
    }
    return DFN_OK;
  }

  int upload_params() {
    for (int i = 0; i < B; i++) h_nets[i] = nets[i];
    DFN_HIP(hipMemcpyAsync(cgnets.p, h_nets, sizeof(CgNet) * B, hipMemcpyHostToDevice, stream));
    DFN_HIP(hipMemcpyAsync(d_h.p, h_h, sizeof(double) * B, hipMemcpyHostToDevice, stream));
    DFN_HIP(hipMemcpyAsync(d_ttry.p, h_ttry, sizeof(double) * B, hipMemcpyHostToDevice, stream));
    DFN_HIP(hipMemcpyAsync(d_active.p, h_active, sizeof(int) * B, hipMemcpyHostToDevice, stream));
    return DFN_OK;
  }

  int readback() {
    DFN_HIP(hipMemcpyAsync(h_status, status.p, sizeof(StatusWord) * B, hipMemcpyDeviceToHost, stream));
    DFN_HIP(hipMemcpyAsync(h_cgres, cg_res.p, sizeof(int) * 2 * kEnsSlots * B, hipMemcpyDeviceToHost, stream));
    DFN_HIP(hipStreamSynchronize(stream));
    return DFN_OK;
  }

  // one iteration of src/solver.f90:59-198 for every network
  // One step of every network with part[i] != 0 (all when part is null); with
  // t_end > 0 a participating network's trial step is clipped so that it lands
  // on t_end.  Networks sitting out keep their state: no launch writes them.
  int step(const std::vector<int> *part = nullptr, double t_end = 0.0) {
    const double SAFETY = 0.9, PGROW = -0.2, PSHRNK = -0.25, ERRCON = (double)1.89e-4f;
    std::vector<double> h(B), hdid(B), hnext(B), err_acc(B, 0.0);
    std::vector<int> todo(B, 1), nrej(B, 0), guarded(B, 0), clipped(B, 0);
    if (part) todo = *part;
    for (int i = 0; i < B; i++) {
      if (!todo[i]) continue;
      const HostProblem &p = hp[i];
      if (t_try[i] > 10.0)  // :64-77
        for (size_t s = 0; s < p.src_tb.size(); s++)
          if ((time[i] < p.src_tb[s]) && (time[i] + t_try[i] > p.src_tb[s])) { t_try[i] = 10.0; guarded[i] = 1; }
      if (t_end > 0.0 && time[i] + t_try[i] >= t_end) { t_try[i] = t_end - time[i]; clipped[i] = 1; }
      nets[i].time = time[i];
    }
    // first attempt (with compute_P(t_try) halving, :81-92)
    std::vector<int> act = todo;
    for (;;) {
      all_active = true;
      for (int i = 0; i < B; i++) {
        h_ttry[i] = t_try[i];
        h_h[i] = t_try[i];
        h_active[i] = act[i];
        if (!act[i]) all_active = false;
      }
      DFN_TRY(upload_params());
      DFN_TRY(run(false));
      all_active = false;
      DFN_TRY(readback());
      bool again = false;
      for (int i = 0; i < B; i++) {
        if (!act[i]) continue;
        const int *r0 = h_cgres + 2 * ((size_t)i * kEnsSlots);
        st[i].cg_iterations += r0[1];
        st[i].cg_solves += 1;
        for (int sl = 0; sl < kEnsSlots; sl++)
          if (r0[2 * sl] == 2) return DFN_ERR_CG_TIMEOUT;  // a cooperative solve lost a message
        if (r0[0]) {  // compute_P(t_try) failed: halve and redo this network
          t_try[i] = t_try[i] / 2.0;
          clipped[i] = 0;
          st[i].cg_halvings++;
          again = true;
        } else {
          act[i] = 0;
        }
      }
      if (!again) break;  // otherwise act holds the networks whose solve failed
    }
    // rkqs (ode_bs.f90:125-174) per network on the attempts just made
    for (int i = 0; i < B; i++) {
      if (!todo[i]) continue;
      h[i] = t_try[i];
      st[i].rhs_evaluations += 1;
    }
    const std::vector<int> stepping = todo;
    bool any = true;
    while (any) {
      any = false;
      for (int i = 0; i < B; i++) {
        if (!todo[i]) { h_active[i] = 0; continue; }
        const int *r = h_cgres + 2 * ((size_t)i * kEnsSlots);
        int fail = 0;
        for (int sl = 1; sl < kEnsSlots; sl++) {
          st[i].cg_iterations += r[2 * sl + 1];
          if (r[2 * sl] == 2) return DFN_ERR_CG_TIMEOUT;
          if (r[2 * sl]) fail = 1;
        }
        st[i].cg_solves += 5;
        st[i].rhs_evaluations += 5;
        if (fail) return DFN_ERR_CG;
        double em;
        std::memcpy(&em, &h_status[i].errmax_bits, sizeof(double));
        const double errmax = em / hp[i].tol;
        if (errmax <= 1.0) {
          err_acc[i] = errmax;
          hdid[i] = h[i];
          hnext[i] = errmax > ERRCON ? SAFETY * h[i] * std::pow(errmax, PGROW) : 5.0 * h[i];
          todo[i] = 0;
          h_active[i] = 0;
          continue;
        }
        st[i].rejected_attempts++;
        nrej[i]++;
        clipped[i] = 0;  // the retried step stops short of t_end
        const double htemp = SAFETY * h[i] * std::pow(errmax, PSHRNK);
        h[i] = std::copysign(std::fmax(std::fabs(htemp), 0.1 * std::fabs(h[i])), h[i]);
        if (0.0 + h[i] == 0.0) return DFN_ERR_UNDERFLOW;
        h_active[i] = 1;
        h_h[i] = h[i];
        any = true;
      }
      if (!any) break;
      DFN_TRY(upload_params());
      DFN_TRY(run(true));
      DFN_TRY(readback());
    }
    // commit (:174-177): the accepted state and compute_P(t_did) of slot 4
    DFN_HIP(hipMemcpyAsync(Y.p, Ynew.p, (size_t)3 * Ntot * sizeof(double), hipMemcpyDeviceToDevice, stream));
    DFN_HIP(hipMemcpyAsync(P.p, cg_vec(4, 10), sizeof(double) * Ntot, hipMemcpyDeviceToDevice, stream));
    DFN_HIP(hipMemcpyAsync(Pdot.p, cg_vec(4, 11), sizeof(double) * Ntot, hipMemcpyDeviceToDevice, stream));
    bool tensile = false;
    for (int i = 0; i < B; i++) {
      if (!stepping[i]) continue;
      time[i] = clipped[i] ? t_end : time[i] + hdid[i];
      st[i].accepted_steps++;
      if (h_status[i].tensile) tensile = true;
    }
    const std::vector<double> t_keep = t_try;
    DFN_TRY(control(hdid, hnext, err_acc, nrej, guarded, stepping, clipped));
    for (int i = 0; i < B; i++) {
      if (!stepping[i]) t_try[i] = t_keep[i];
      if (t_try[i] > 24 * 60 * 60) t_try[i] = 24 * 60 * 60;
    }
    return tensile ? DFN_ERR_TENSILE : DFN_OK;
  }
};

}  // namespace dfn

using dfn::EnsEngine;

struct dfn_ensemble_s {
  EnsEngine e;
};

#define API extern "C" __attribute__((visibility("default")))
#include "api_guard.hpp"

API int dfn_ensemble_create(const dfn_problem *problems, int32_t nnet, dfn_ensemble *out) {
  DFN_API_BEGIN
  if (!out || !problems || nnet < 1) return DFN_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return DFN_ERR_HIP;
  dfn_ensemble h = new (std::nothrow) dfn_ensemble_s;
  if (!h) return DFN_ERR_NOMEM;
  const int st = h->e.init(problems, nnet);
  if (st != DFN_OK) { delete h; return st; }
  *out = h;
  return DFN_OK;
  DFN_API_END
}

API int dfn_ensemble_destroy(dfn_ensemble h) {
  DFN_API_BEGIN
  if (!h) return DFN_ERR_ARG;
  delete h;
  return DFN_OK;
  DFN_API_END
}

API int dfn_ensemble_build_hmat(dfn_ensemble h) {
  DFN_API_BEGIN
  if (!h) return DFN_ERR_ARG;
  return h->e.build_hmat();
  DFN_API_END
}

API int dfn_ensemble_set_precision(dfn_ensemble h, int32_t prec) {
  DFN_API_BEGIN
  if (!h || (prec < DFN_PREC_FP64 || prec > DFN_PREC_MIXED_SPLIT)) return DFN_ERR_ARG;
  h->e.precision = prec;
  if (prec != DFN_PREC_FP64 && h->e.ops32_kind != prec) return h->e.layout_reduced();
  h->e.drop_graphs();
  return DFN_OK;
  DFN_API_END
}

API int dfn_ensemble_set_graph(dfn_ensemble h, int32_t use_graph) {
  DFN_API_BEGIN
  if (!h) return DFN_ERR_ARG;
  h->e.use_graph = use_graph != 0 && !h->e.coop;  // cooperative launches are not captured
  return DFN_OK;
  DFN_API_END
}

API int dfn_ensemble_solve(dfn_ensemble h, int64_t nsteps) {
  DFN_API_BEGIN
  if (!h || nsteps < 0) return DFN_ERR_ARG;
  if (!h->e.active_ops()) return DFN_ERR_STATE;
  for (int64_t s = 0; s < nsteps; s++) DFN_TRY(h->e.step());
  return DFN_OK;
  DFN_API_END
}

API int dfn_ensemble_solve_until(dfn_ensemble h, double t_end, int64_t max_steps) {
  DFN_API_BEGIN
  if (!h || !(t_end > 0.0) || max_steps < 0) return DFN_ERR_ARG;
  EnsEngine &e = h->e;
  if (!e.active_ops()) return DFN_ERR_STATE;
  for (int64_t s = 0; s < max_steps; s++) {
    std::vector<int> part(e.B, 0);
    bool any = false;
    for (int i = 0; i < e.B; i++)
      if (e.time[i] < t_end) { part[i] = 1; any = true; }
    if (!any) return DFN_OK;
    DFN_TRY(e.step(&part, t_end));
  }
  for (int i = 0; i < e.B; i++)
    if (e.time[i] < t_end) return DFN_ERR_STATE;  // max_steps ran out first
  return DFN_OK;
  DFN_API_END
}

API int dfn_ensemble_get_state(dfn_ensemble h, int32_t netw, double *Yh, double *Ph, double *Pdh, double *time) {
  DFN_API_BEGIN
  if (!h || netw < 0 || netw >= h->e.B) return DFN_ERR_ARG;
  EnsEngine &e = h->e;
  const size_t N = e.hp[netw].N, o = e.base[netw], T = e.Ntot;
  DFN_HIP(hipStreamSynchronize(e.stream));
  if (Yh)
    for (int c = 0; c < 3; c++)
      DFN_HIP(hipMemcpy(Yh + c * N, e.Y.as<double>() + c * T + o, N * sizeof(double), hipMemcpyDeviceToHost));
  if (Ph) DFN_HIP(hipMemcpy(Ph, e.P.as<double>() + o, N * sizeof(double), hipMemcpyDeviceToHost));
  if (Pdh) DFN_HIP(hipMemcpy(Pdh, e.Pdot.as<double>() + o, N * sizeof(double), hipMemcpyDeviceToHost));
  if (time) *time = e.time[netw];
  return DFN_OK;
  DFN_API_END
}

API int dfn_ensemble_get_stats(dfn_ensemble h, int32_t netw, dfn_run_stats *s) {
  DFN_API_BEGIN
  if (!h || !s || netw < 0 || netw >= h->e.B) return DFN_ERR_ARG;
  EnsEngine &e = h->e;
  *s = e.st[netw];
  s->time = e.time[netw];
  s->t_try = e.t_try[netw];
  const size_t N = e.hp[netw].N;
  std::vector<double> V(N);
  DFN_HIP(hipStreamSynchronize(e.stream));
  DFN_HIP(hipMemcpy(V.data(), e.Y.as<double>() + e.base[netw], N * sizeof(double), hipMemcpyDeviceToHost));
  s->max_V = *std::max_element(V.begin(), V.end());
  return DFN_OK;
  DFN_API_END
}

API int dfn_ensemble_set_controller(dfn_ensemble h, int32_t mode, dfn_mlp m) {
  DFN_API_BEGIN
  if (!h || mode < DFN_STEP_RKQS || mode > DFN_STEP_LEARNED) return DFN_ERR_ARG;
  if (mode == DFN_STEP_LEARNED && !m) return DFN_ERR_ARG;
  EnsEngine &e = h->e;
  e.controller = mode;
  e.mlp = mode == DFN_STEP_LEARNED ? dfn::mlp_of(m) : nullptr;
  std::fill(e.fallback.begin(), e.fallback.end(), 0);
  std::fill(e.learned_last.begin(), e.learned_last.end(), 0);
  return DFN_OK;
  DFN_API_END
}

API int dfn_ensemble_set_recording(dfn_ensemble h, int32_t on) {
  DFN_API_BEGIN
  if (!h) return DFN_ERR_ARG;
  EnsEngine &e = h->e;
  e.recording = on != 0;
  std::fill(e.has_pend.begin(), e.has_pend.end(), 0);
  return DFN_OK;
  DFN_API_END
}

API int dfn_ensemble_take_records(dfn_ensemble h, float *X, float *T, int64_t cap, int64_t *count) {
  DFN_API_BEGIN
  if (!h || !count || cap < 0) return DFN_ERR_ARG;
  EnsEngine &e = h->e;
  const int64_t avail = (int64_t)e.rec_T.size();
  if (!X || !T) { *count = avail; return DFN_OK; }
  const int64_t n = std::min(cap, avail);
  std::copy(e.rec_X.begin(), e.rec_X.begin() + n * DFN_MLP_INPUTS, X);
  std::copy(e.rec_T.begin(), e.rec_T.begin() + n, T);
  e.rec_X.erase(e.rec_X.begin(), e.rec_X.begin() + n * DFN_MLP_INPUTS);
  e.rec_T.erase(e.rec_T.begin(), e.rec_T.begin() + n);
  *count = n;
  return DFN_OK;
  DFN_API_END
}

API int dfn_ensemble_kernel_timing(dfn_ensemble h, int32_t on) {
  DFN_API_BEGIN
  if (!h) return DFN_ERR_ARG;
  EnsEngine &e = h->e;
  if (on) {
    if (!e.kt_on) e.kt_graph = e.use_graph;
    e.use_graph = false;  // events are recorded around launches issued one by one
    e.kt_on = true;
    e.kt_used = 0;
    e.kt_ms = 0.0;
    e.kt_count = 0;
  } else if (e.kt_on) {
    DFN_TRY(e.kt_collect());
    e.kt_on = false;
    e.use_graph = e.kt_graph && !e.coop;
  }
  return DFN_OK;
  DFN_API_END
}

API int dfn_ensemble_kernel_timing_read(dfn_ensemble h, double *ms_mean, int64_t *count, double *bytes) {
  DFN_API_BEGIN
  if (!h) return DFN_ERR_ARG;
  EnsEngine &e = h->e;
  DFN_TRY(e.kt_collect());
  if (ms_mean) *ms_mean = e.kt_count ? e.kt_ms / (double)e.kt_count : 0.0;
  if (count) *count = e.kt_count;
  if (bytes) *bytes = e.active_ops() ? e.active_ops()->dom_bytes() : 0.0;
  return DFN_OK;
  DFN_API_END
}

API int dfn_ensemble_info(dfn_ensemble h, int64_t *info) {
  DFN_API_BEGIN
  if (!h || !info) return DFN_ERR_ARG;
  EnsEngine &e = h->e;
  const dfn::DevOps *d = e.active_ops();
  info[0] = e.B;
  info[1] = e.Ntot;
  info[2] = e.max_n;
  info[3] = d ? d->S_total : 0;
  info[4] = d ? (int64_t)d->alg_bytes : 0;
  info[5] = e.use_graph ? 1 : 0;
  info[6] = d && d->fused ? 1 : 0;
  info[7] = d ? d->kc : 0;
  info[8] = d && d->vpair ? 1 : 0;
  return DFN_OK;
  DFN_API_END
}
