// Device-side data layouts shared by the kernels and the host engine.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace dfn {

constexpr int kMaxRank = 32;          // largest low-rank block rank the phase-1 kernel keeps in registers
constexpr int kPanelMaxCols = 2048;   // columns per phase-2 panel task (w staged in 16 KB of LDS)
constexpr int kSplitMaxCols = 1024;   // columns per task of split near-field panels (w in FP64 and FP32: 12 KB)
constexpr int kSplitBytes = 5;        // bytes per split near-field entry (FP32 value, 8-bit residual)
constexpr int kVChunkRows = 512;      // rows of V per phase-1 wave task
constexpr int kXPad = 4;              // zero entries after x_h in w: phase 1 reads whole 16-byte vectors
constexpr int kCgMaxIter = 200;       // compute_P gives up after 200 iterations (fluid_diffusion.f90:120)

// Phase 1: out[j] = scale * sum_{i<len} V[row0+i, j] * w[xcol+row0+i], j < k.
// FP32 storage holds V / scale with a power-of-two scale per block.
struct VTask {
  int64_t voff;   // offset of this chunk's V, stored in strips of 64 x (16-byte vector) rows
  double scale;
  int32_t row0;   // first row of this chunk
  int32_t len;    // rows in this chunk
  int32_t k;      // rank
  int32_t xcol;   // H index of the block's first column
  int32_t zp;     // output offset in zpart (blocks split in several chunks)
  int32_t wout;   // >= 0: single-chunk block, z written straight to w[wout + j]
  int32_t net;    // network of the block (ensembles; 0 for one network)
};

// One wave of phase 1: task a alone (b < 0), or tasks a and b, of the same rank and
// network and each a single strip of at most 32 row vectors, on the wave's two halves.
struct VItem {
  int32_t a, b;
};

// z of one low-rank block: w[zout + j] = sum_t zpart[zp_first + t*k + j]
struct ZRed {
  int32_t zp_first, ntask, k, zout;
};

// Phase 2: ypart[yp + r] = sum_j P[r, j] * w[widx[j]] (* cscale[j] in FP32 storage,
// where the first nd columns are near-field entries summed with compensation)
// over a row-major m x L panel.
struct PTask {
  int64_t poff;
  int64_t widx;
  int32_t m;
  int32_t L;
  int32_t yp;
  int32_t nd;
  int32_t net;  // network of the panel's rows (ensembles; 0 for one network)
  int32_t pad;
};

struct RhsParams {
  const double *a, *b, *Dc, *f0, *V0, *nld, *sld;
  const uint8_t *fake;
  double eta_rad;  // mu / (2 cs), radiation damping
  int mode3;
};

struct StatusWord {
  unsigned long long errmax_bits;  // max |yerr/yscal| as the bits of a non-negative double
  int tensile;
  int pad;
};

// Newton-V form (an extension with no reference counterpart): the state holds
// slip instead of slip rate, and each evaluation solves the stress balance
// tau = eta V - (sigma + P) a asinh(V / (2 V0) exp((f0 + b psi) / a)) for V,
// with tau = tau_ref + shear_loading_dot t + K slip and
// sigma = sig_ref + normal_loading_dot t (+ K_N slip in Mode II), t measured
// from the switch to this form.  on = 0: the reference form.
struct NvArgs {
  const double *tau_ref, *sig_ref;
  double *Vg;  // last solved slip rate per element (the next solve's first guess)
  double t;    // time of this evaluation since the switch
  int on;
};

struct StageArgs {
  int n;
  RhsParams p;
  NvArgs nv;
  const double *y;      // step-start state (3N)
  const double *ytemp;  // this stage's input state (3N)
  const double *sT, *sN, *Pt, *Pd;
  double *k[6];
  double *ynext;        // next stage input, or the 5th-order solution at stage 6
  double *yscal;
  double h, t_try, tol;
  StatusWord *status;
};

struct CgSys {
  double dt;
  double *diag, *A1, *A2, *A3, *b, *x, *r, *p, *Ap;
  double *Pt, *Pd;
  int *result;  // [ier, iterations]
  int net;      // network of the system (multi-network launches)
  int pad;
};

// One network of a multi-network compute_P launch: its elements are [base,
// base + n) of the concatenated arrays (permeability_x at xbase, n + 1 entries),
// its sources and crossings the given ranges of the concatenated lists (node ids
// and element ids stay network-local), and its own fluid constants and time.
struct CgNet {
  int n, base, xbase, src0, nsrc, cross0, ncross, pad;
  double cfac, time, tol, por, beta;
};

struct CgArgs {
  int n;
  const CgSys *sys;
  const double *perm_x;       // permeability_x (0-based: perm_x[k] = permeability_x(k+1))
  const double *perm_x_diag;  // permeability_x in effect for compute_diag
  const double *perm;         // permeability for the crossing terms of the stencil
  const double *perm_diag;    // permeability for the crossing terms of compute_diag
  const double *ds;
  const double *P;            // step-start pressure
  double cfac;                // porosity * dyn_viscosity * (fluid_comp + rock_comp)
  double time, tol, por, beta;
  int nsrc;
  const int64_t *src_elem;
  const double *src_Q, *src_tb, *src_te;
  int ncross;
  const int64_t *cross;
  const CgNet *nets;   // null: one network described by the fields above
  const int *active;   // multi-network launches: systems of networks with active[net] == 0 return at once
  int late_part;       // test hook of the cooperative kernel: workgroup late_part - 1 of every system
                       // reports that a message never arrived (0: off)
  // multi-network cooperative launches (may be null): per workgroup {system, part,
  // the system's workgroup count, the system's first workgroup in the launch}
  const int4 *wg_map;
};

// Crossing rows of the fluid-diffusion stencil: entries (crossing, slot) of each
// element, in crossing order (slot 0: element node1-1, 1: node1, 2: node2-1, 3: node2).
struct CrossRows {
  const int *start;  // [n+1]
  const int *ent;    // [2*entries] (crossing index, slot)
};

constexpr int kCgLdsThreads = 512;          // threads of the one-workgroup-per-system CG
constexpr int kCgLdsMax = kCgLdsThreads * 8;  // largest system it takes (elements)
constexpr int kCgEPT = 8;                   // elements per thread in the cooperative CG
constexpr int kCgSpan = kCgEPT * 256;       // elements per workgroup
constexpr int kCgCoopMaxEntries = 896;      // crossing entries per workgroup span the cooperative CG stages in LDS
constexpr int kCgRecordWords = 8;           // message words per workgroup of the cooperative CG
constexpr unsigned kCgMaxSeq = 1u << 24;    // launch sequence numbers of the cooperative CG are below this

struct PermArgs {
  int n, coupling;
  double *perm, *perm_x;
  const double *perm_star, *kmin, *sigmaN, *P, *Snk;
};

// launchers (kernels.hip)
// xf: FP32 copy of x_h for FP32 operators (n + 4 floats), or nullptr
void launch_gather(const double *x, const int *perm, double *xh, int n, hipStream_t s, float *xf = nullptr);
// storage and arithmetic of an operator's factors and panels (mode of launch_vtx and
// launch_panel): FP64; FP32 with FP32 products and sums (compensated on near-field
// columns); FP32 storage widened into FP64 products and sums
// kStoreF64NT: FP64 factors read with non-temporal loads (phase 1 only)
enum { kStoreF64 = 0, kStoreF32 = 1, kStoreF32Acc64 = 2, kStoreF64NT = 3 };
// waves: tasks (waves) per workgroup, 2, 4 or 8; xf: the FP32 copy (kStoreF32);
// active (ensembles, may be null): tasks of networks with active[net] == 0 return at once;
// kmax: the largest rank of any task (at most kMaxRank)
// items (may be null): the waves' task lists (nitems of them) in place of one task per wave
void launch_vtx(const VTask *t, int nt, const void *vdata, int mode, double *w, double *zpart, hipStream_t s,
                int waves = 4, const float *xf = nullptr, const int *active = nullptr, int kmax = kMaxRank,
                const VItem *items = nullptr, int nitems = 0);
// the rank ceiling of the phase-1 kernel instance that serves tasks of rank <= kmax
int vtx_rank_ceiling(int kmax);
void launch_zreduce(const ZRed *zr, int nb, const double *zpart, double *w, hipStream_t s);
// phase 1 over FP32 factors with FP64 sums together with the FP64 near-field panels
// (pt: the FP64 panel set of a mixed-precision layout) in one launch; rows and
// active as for launch_panel
void launch_vtx_near(const VTask *vt, int nvt, const float *vdata, const PTask *pt, int npt, const double *pdata,
                     const int *widx, double *w, double *zpart, double *ypart, hipStream_t s, int rows = 2,
                     const int *active = nullptr, int kmax = kMaxRank);
// near-field panels stored as FP32 values (hi) and 8-bit residuals (lo, same entry
// offsets; cscale at a task's first column the residual's scale); rows and active
// as for launch_panel
void launch_panel_split(const PTask *t, int nt, const float *hi, const signed char *lo, const int *widx,
                        const float *cscale, const double *w, double *ypart, hipStream_t s, int rows = 2,
                        const int *active = nullptr);
// rows: panel rows per wave step, 2 or 4; active as for launch_vtx
void launch_panel(const PTask *t, int nt, const void *pdata, int mode, const int *widx, const float *cscale,
                  const double *w, double *ypart, hipStream_t s, int rows = 2, const int *active = nullptr);
// Multi-RHS stress evaluation (FP64 layout): X holds nrhs vectors of n (vector
// after vector, original order), y{T,N} likewise; R is nrhs rounded up to a multiple
// of 64; W ((w size) x R), zpart ((zpart size) x R), ypart ((ypart size) x R) and Yh
// (nops n x R) are work arrays, row-major with row stride R; ev (optional) gets
// events recorded before and after the phase-2 kernel.
struct MrhsLayout {
  const VTask *vt;
  const double *vdata;
  const ZRed *zr;
  const PTask *pt;
  const double *pdata;
  const int *widx, *fin_first, *fin_idx, *perm;
  int nvt, nzr, npt, n, nops;
};
void launch_mrhs(const MrhsLayout &d, const double *X, int nrhs, int R, double *W, double *zpart, double *ypart,
                 double *Yh, double scale, double *yT, double *yN, hipStream_t s, hipEvent_t *ev);
void launch_finalize(const int *fin_first, const int *fin_idx, const int *perm, const double *ypart, int n, int nops,
                     double scale, double *yT, double *yN, hipStream_t s);
// dense direct kernels: node columns per tile, number of tiles along the nodes,
// and the partial-sum buffer each launch needs (doubles): direct_chunks(n) * n
// for Mode III, twice that for Mode II
constexpr int kDirectCols = 64;
int direct_chunks(int n);
void launch_direct3(const double *node, const double *slip, double *part, int n, double scale, double *sT,
                    hipStream_t s);
void launch_direct2(const double *node, const double *elem, const double *nrm, const double *slip, double *part, int n,
                    double scale, double *sT, double *sN, hipStream_t s);
void launch_stage(int stage, const StageArgs &a, hipStream_t s);
// Newton-V form: from the reference-form state Y = [V | psi | sigma_N] and P, the
// balance's reference stress and normal stress, first guesses, and Y's slip and
// third blocks set to zero (nv_enter); the reference-form state of a Newton-V
// state whose stress transfer K slip is in sT / sN (nv_state, t since the switch)
void launch_nv_enter(const RhsParams &p, double *Y, const double *P, int n, double *tau_ref, double *sig_ref,
                     double *Vg, hipStream_t s);
void launch_nv_state(const RhsParams &p, const NvArgs &nv, const double *Y, const double *P, const double *sT,
                     const double *sN, int n, double *out, hipStream_t s);

// Ensemble of networks concatenated element-wise (network of element k: net[k]):
// the stage kernels with per-network step sizes, masks and error norms.
struct EnsStageArgs {
  StageArgs a;              // n, params, buffers over all networks; a.h / a.t_try / a.tol / a.status unused
  const int *net;           // [n]
  const int *active;        // [nnet] 1: the network takes part in this attempt
  const double *h, *t_try, *tol, *eta;  // [nnet]
  StatusWord *status;       // [nnet]
};
void launch_ens_stage(int stage, const EnsStageArgs &a, hipStream_t s);
void launch_ens_restart_stage2(const EnsStageArgs &a, hipStream_t s);
void launch_ens_clear_status(StatusWord *st, const int *active, int nnet, hipStream_t s);
// event monitor: slip += (V_old + V_new) h / 2 and per-fault max |V_new| (cleared first)
void launch_save_output(const double *Y, const double *P, const double *sT, const double *perm, double *buf, int n,
                        int64_t ld, int64_t col, hipStream_t s);
void launch_monitor(const double *Vold, const double *Vnew, double h, const int *fault_of, int n, double *slip,
                    unsigned long long *fvmax, int nfault, hipStream_t s);
// per network, over its real elements: max |V|, mean log10 |V|, fraction with
// |V| > 1e-3 m/s (out[3 b + 0..2]), reduced in a fixed order
void launch_ens_vstats(const double *V, const uint8_t *fake, const CgNet *nets, int nnet, double *out, hipStream_t s);
// dt of every (network, slot) system of an ensemble compute_P launch: slot 0 and,
// when with_stages, slots 1..5 (x + A_i h with x = 0), from t_try (slot 0) and h
void launch_ens_cg_dt(CgSys *sys, int nnet, int nslot, const double *t_try, const double *h, int first_slot,
                      hipStream_t s);
void launch_cg_lds_multi(const CgArgs &a, const CrossRows &xr, int nsys, int max_n, hipStream_t s);
void launch_restart_stage2(const double *y, const double *k1, double h, int n3, double *yt, hipStream_t s);
void launch_cg(const CgArgs &a, int nsys, hipStream_t s);
bool launch_cg_lds(const CgArgs &a, const CrossRows &xr, int nsys, hipStream_t s);
// cooperative multi-workgroup CG; cap = the most crossing entries of any span of
// kCgSpan rows; max_blocks = cg_coop_capacity(cap) on the current device; seq in
// [1, 2^24) differs from the seq of every earlier launch on the same records and
// direction buffers since they were last cleared; records holds kCgRecordWords
// words per workgroup of each system.  Returns false if a span holds more than
// kCgCoopMaxEntries or one system's workgroups cannot be co-resident.  A
// system's result[0] is 2 when a workgroup waited in vain for a message.
bool launch_cg_coop(const CgArgs &a, const CrossRows &xr, int nsys, int cap, int max_blocks, unsigned seq,
                    unsigned long long *records, hipStream_t s);
// Multi-network cooperative launch of nwg workgroups described by a.wg_map (all
// co-resident; nwg <= cg_coop_capacity(cap)); false if the launch fails.
bool launch_cg_coop_map(const CgArgs &a, const CrossRows &xr, int nwg, int cap, unsigned seq,
                        unsigned long long *records, hipStream_t s);
// workgroups of cg_coop_kernel that can be resident together (0: none / unsupported)
int cg_coop_capacity(int cap);
void launch_perm_update(const PermArgs &a, hipStream_t s);
void launch_stream_copy(const void *a, void *b, int64_t bytes, hipStream_t s);

}  // namespace dfn
