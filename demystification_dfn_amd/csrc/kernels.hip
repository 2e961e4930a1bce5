// gfx950 kernels of the per-RK-stage hot path.
//
//   * stress-transfer H-matvec in two phases over device-resident block lists:
//       phase 1  z_b = V_b^T x      one wave per (low-rank block, row chunk)
//       phase 2  y_R = [F | U] w    one workgroup per leaf-row panel (row-major,
//                                   w = [x_h | z] gathered through LDS)
//       finalize y(perm) = scale * sum of panel partials (fixed order, no atomics)
//   * directIII / directII on the fly (dense dislocation kernels)
//   * rate-state right-hand side fused with the Cash-Karp stage combinations and
//     the max-norm error reduction
//   * compute_P: row-scaled CG, several time steps solved side by side
//
// Kernels whose results feed the ODE state are compiled without floating-point
// contraction so that every +,-,*,/ rounds as in the reference's expressions.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dfn_device.hpp"

namespace dfn {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Wave reductions through DPP lane moves, for call sites where all 64 lanes are
// active: xor-1 and xor-2 quad permutes, the half-row and row mirrors give every
// lane its 16-lane row total (a + b and b + a are the same double, so the lanes
// of a row agree bit for bit); the four row totals are then read from lanes 0,
// 16, 32, 48 and added in that order.  Every lane returns the same value.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double lane_value(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}

constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140;

__device__ __forceinline__ void wave_sum_max_full(double &sum, double &mx) {
  sum += dpp_move<kDppXor1>(sum);
  mx = fmax(mx, dpp_move<kDppXor1>(mx));
  sum += dpp_move<kDppXor2>(sum);
  mx = fmax(mx, dpp_move<kDppXor2>(mx));
  sum += dpp_move<kDppHalfMirror>(sum);
  mx = fmax(mx, dpp_move<kDppHalfMirror>(mx));
  sum += dpp_move<kDppMirror>(sum);
  mx = fmax(mx, dpp_move<kDppMirror>(mx));
  sum = (lane_value(sum, 0) + lane_value(sum, 16)) + (lane_value(sum, 32) + lane_value(sum, 48));
  mx = fmax(fmax(lane_value(mx, 0), lane_value(mx, 16)), fmax(lane_value(mx, 32), lane_value(mx, 48)));
}

__device__ __forceinline__ double wave_sum_full(double sum) {
  sum += dpp_move<kDppXor1>(sum);
  sum += dpp_move<kDppXor2>(sum);
  sum += dpp_move<kDppHalfMirror>(sum);
  sum += dpp_move<kDppMirror>(sum);
  return (lane_value(sum, 0) + lane_value(sum, 16)) + (lane_value(sum, 32) + lane_value(sum, 48));
}

template <int CTRL>
__device__ __forceinline__ float dpp_move_f(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}

__device__ __forceinline__ float wave_sum_full(float sum) {
  sum += dpp_move_f<kDppXor1>(sum);
  sum += dpp_move_f<kDppXor2>(sum);
  sum += dpp_move_f<kDppHalfMirror>(sum);
  sum += dpp_move_f<kDppMirror>(sum);
  const auto at = [](float v, int l) { return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l)); };
  return (at(sum, 0) + at(sum, 16)) + (at(sum, 32) + at(sum, 48));
}

// Reduce-scatter of N per-lane values over the 64 lanes of a wave (all lanes
// active), one step per lane bit: each step sums every value over a lane pair
// and leaves the pair's lanes half of the values each (the lane with the bit clear
// keeps the first half), until a lane holds one value; later steps are plain pair
// sums.  Bits 5 and 4 go through v_permlane32_swap / v_permlane16_swap, which
// hand each lane its partner's copy of the value the lane keeps (no select);
// bits 3..0 through DPP row-mirror, half-row-mirror, xor-2 and xor-1 moves.  A
// K-column reduction costs about 2K value moves in all, against 6K to 8K for K
// whole-wave reductions.  Every pair sum is a + b on both lanes, so lanes that
// hold the same value hold it bit for bit.
__device__ __forceinline__ double swap32_sum(double a, double b) {
  // lanes 0-31: a(l) + a(l + 32); lanes 32-63: b(l - 32) + b(l)
  const auto lo = __builtin_amdgcn_permlane32_swap(__double2loint(a), __double2loint(b), false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap(__double2hiint(a), __double2hiint(b), false, false);
  return __hiloint2double(hi[0], lo[0]) + __hiloint2double(hi[1], lo[1]);
}
__device__ __forceinline__ float swap32_sum(float a, float b) {
  // the two results are copied to scalars first: __builtin_bit_cast applied to an
  // element of the returned vector read element 0 for both
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  const unsigned r0 = r[0], r1 = r[1];
  return __uint_as_float(r0) + __uint_as_float(r1);
}
__device__ __forceinline__ double swap16_sum(double a, double b) {
  // even 16-lane rows: a(l) + a(l + 16); odd rows: b(l - 16) + b(l)
  const auto lo = __builtin_amdgcn_permlane16_swap(__double2loint(a), __double2loint(b), false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap(__double2hiint(a), __double2hiint(b), false, false);
  return __hiloint2double(hi[0], lo[0]) + __hiloint2double(hi[1], lo[1]);
}
__device__ __forceinline__ float swap16_sum(float a, float b) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  const unsigned r0 = r[0], r1 = r[1];
  return __uint_as_float(r0) + __uint_as_float(r1);
}
template <int CTRL>
__device__ __forceinline__ double lane_move(double v) { return dpp_move<CTRL>(v); }
template <int CTRL>
__device__ __forceinline__ float lane_move(float v) { return dpp_move_f<CTRL>(v); }

template <int STEP, int N, typename T>
__device__ __forceinline__ void rs_step(const T (&in)[N], T (&out)[(N + 1) / 2], int lane) {
  constexpr int H = (N + 1) / 2;
  if constexpr (STEP >= 2 && N == 1) {  // one value left: a plain pair sum
    constexpr int C = STEP == 2 ? kDppMirror : STEP == 3 ? kDppHalfMirror : STEP == 4 ? kDppXor2 : kDppXor1;
    out[0] = in[0] + lane_move<C>(in[0]);
  } else {
#pragma unroll
    for (int j = 0; j < H; j++) {
      const T a = in[j], b = j + H < N ? in[j + H < N ? j + H : 0] : (T)0;
      if constexpr (STEP == 0) {
        out[j] = swap32_sum(a, b);
      } else if constexpr (STEP == 1) {
        out[j] = swap16_sum(a, b);
      } else {
        constexpr int C = STEP == 2 ? kDppMirror : STEP == 3 ? kDppHalfMirror : STEP == 4 ? kDppXor2 : kDppXor1;
        const bool up = (lane >> (5 - STEP)) & 1;
        out[j] = (up ? b : a) + lane_move<C>(up ? a : b);
      }
    }
  }
}

// the reduce-scatter of acc[0..K) (K <= 64); returns the lane's value and sets col
// to the column it totals, or -1 (a lane of a pair whose other lane holds it).
// HALVES: the two 32-lane halves hold separate sums (lane bit 5 takes no part), and
// each half ends with its own totals.
template <int K, bool HALVES = false, typename T>
__device__ __forceinline__ T wave_reduce_scatter(const T (&acc)[K], int lane, int &col) {
  constexpr int N1 = HALVES ? K : (K + 1) / 2, N2 = (N1 + 1) / 2, N3 = (N2 + 1) / 2, N4 = (N3 + 1) / 2;
  constexpr int N5 = (N4 + 1) / 2, N6 = (N5 + 1) / 2;
  T a1[N1], a2[N2], a3[N3], a4[N4], a5[N5], a6[N6];
  if constexpr (HALVES) {
#pragma unroll
    for (int j = 0; j < K; j++) a1[j] = acc[j];
  } else {
    rs_step<0>(acc, a1, lane);
  }
  rs_step<1>(a1, a2, lane);
  rs_step<2>(a2, a3, lane);
  rs_step<3>(a3, a4, lane);
  rs_step<4>(a4, a5, lane);
  rs_step<5>(a5, a6, lane);
  // the column, from the last step back: a splitting step adds bit * (its output
  // count) to the index, which must stay below its input count; a plain step keeps
  // the lane with the bit clear
  const int nin[6] = {K, N1, N2, N3, N4, N5}, nout[6] = {N1, N2, N3, N4, N5, N6};
  int idx = 0;
  bool ok = true;
#pragma unroll
  for (int s = 5; s >= (HALVES ? 1 : 0); s--) {
    const bool bit = (lane >> (5 - s)) & 1;
    if (s < 2 || nin[s] > 1) {
      idx += bit ? nout[s] : 0;
      ok = ok && idx < nin[s];
    } else {
      ok = ok && !bit;
    }
  }
  col = ok ? idx : -1;
  return a6[0];
}

// the R row sums of a panel's wave step, reduced across the wave and stored by the
// lanes that hold them (rows from nrows on are not stored).  The pairing of lanes
// is the same for every R, so a row's sum does not depend on R.
template <int R>
__device__ __forceinline__ void store_row_sums(const double (&a)[R], int lane, double *__restrict__ out, int nrows) {
  int col;
  const double t = wave_reduce_scatter<R>(a, lane, col);
  if (col >= 0 && col < nrows) out[col] = t;
}

// ---------------------------------------------------------------------------
// H-matvec
// ---------------------------------------------------------------------------
// x_h = x in H order; with FP32 operators also its FP32 copy, which phase 1 reads
// (the FP32 body rounds each x_h entry to FP32 before use, so the copy holds the
// same values and is read with one 16-byte load per lane instead of four 8-byte ones)
__global__ void __launch_bounds__(256) gather_perm_kernel(const double *__restrict__ x, const int *__restrict__ perm,
                                                          double *__restrict__ xh, float *__restrict__ xf, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double v = x[perm[i]];
    xh[i] = v;
    if (xf) xf[i] = (float)v;
  }
}

// A vector of V factors; NT: a non-temporal load (an autotuner candidate for FP64
// factors: the values read are the same either way)
template <bool NT, typename VEC>
__device__ __forceinline__ VEC load_v(const VEC *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

// Phase 1: one wave per task.  The body is specialised on the rank K (a
// wave-uniform switch), so the K column loads of a row strip are all in flight
// together; each lane takes a 16-byte vector of consecutive rows (2 FP64 / 4 FP32),
// K partial sums stay in registers and are reduced across the wave with DPP lane
// moves.  The task's V is stored strip by strip (64 vectors of rows, column after
// column; the last strip's height padded with zero rows to a whole vector), so the
// K loads of one step read one contiguous run.
template <int K, typename TV, bool ACC64, bool GROUPS, bool NT = false>
__device__ __forceinline__ void vtx_body(const VTask &task, const TV *__restrict__ vdata, const double *__restrict__ w,
                                         const float *__restrict__ xf, double *__restrict__ zpart,
                                         double *__restrict__ wz, int lane) {
  constexpr int kVec = 16 / (int)sizeof(TV);
  typedef TV tvec __attribute__((ext_vector_type(kVec)));
  // FP64 storage accumulates in FP64; FP32 storage in FP32 (TV), or with ACC64 in
  // FP64 from the FP64 input (the stored entries widened exactly)
  typedef typename std::conditional<ACC64, double, TV>::type TA;
  TA acc[K];
#pragma unroll
  for (int j = 0; j < K; j++) acc[j] = (TA)0;
  constexpr int kStrip = 64 * kVec;
  const TV *V = vdata + task.voff;
  const double *x = w + task.xcol + task.row0;
  for (int s0 = 0; s0 < task.len; s0 += kStrip) {
    const int i = s0 + lane * kVec;
    const int h = min(kStrip, task.len - s0);
    const int hp = (h + kVec - 1) / kVec * kVec;  // the strip's column stride
    if (i >= task.len) break;
    const TV *Vs = V + (int64_t)s0 * K + lane * kVec;  // the strips before hold kStrip rows each
    TA xi[kVec];
    if constexpr (sizeof(TV) == 4 && !ACC64) {
      // four consecutive entries of the FP32 copy (4-byte aligned; the copy is
      // padded past n); past the chunk's rows they meet the zero rows of V
      float4 x4;
      __builtin_memcpy(&x4, xf + task.xcol + task.row0 + i, sizeof(x4));
      xi[0] = x4.x; xi[1] = x4.y; xi[2] = x4.z; xi[3] = x4.w;
    } else {
      // kVec consecutive entries of x_h, read as 16-byte vectors (8-byte aligned);
      // past the chunk's rows they meet the zero rows of V, and x_h is followed by
      // zero entries in w, so every entry read is a finite value times zero
      double xv[kVec];
      __builtin_memcpy(xv, x + i, sizeof(xv));
#pragma unroll
      for (int u = 0; u < kVec; u++) xi[u] = (TA)xv[u];
    }
    // FP32 factors with FP64 sums (GROUPS): columns in groups of 10, so that 10
    // vectors of V are live at a time (96 VGPRs at ranks up to 20: 5 waves per SIMD);
    // each column's sum is the same whatever the grouping
    constexpr int G = (GROUPS && ACC64 && sizeof(TV) == 4 && K > 10) ? 10 : K;
#pragma unroll
    for (int j0 = 0; j0 < K; j0 += G) {
      tvec v[G];
#pragma unroll
      for (int j = 0; j < G; j++)
        if (j0 + j < K) v[j] = load_v<NT>(reinterpret_cast<const tvec *>(Vs + (j0 + j) * hp));
#pragma unroll
      for (int j = 0; j < G; j++)
        if (j0 + j < K)
#pragma unroll
          for (int u = 0; u < kVec; u++) acc[j0 + j] += (TA)v[j][u] * xi[u];
      if constexpr (G < K) __builtin_amdgcn_sched_barrier(0);  // keep the next group's loads after these sums
    }
  }
  // all 64 lanes are active here: one lane per column ends with its total
  int col;
  const TA mine = wave_reduce_scatter<K>(acc, lane, col);
  if (col >= 0) {
    const double v = (double)mine * task.scale;
    if (task.wout >= 0) wz[task.wout + col] = v;
    else zpart[task.zp + col] = v;
  }
}

// one phase-1 task on one wave: a wave-uniform switch on the task's rank.  KMAX is
// the largest rank of the layout rounded up to 16, 20 or 32: the register budget of
// the whole kernel is that of its largest case, so a layout whose ranks stay at or
// below 20 keeps more waves resident (FP64 storage 88 instead of 135 VGPRs)
template <typename TV, bool ACC64, int KMAX, bool GROUPS, bool NT = false>
__device__ __forceinline__ void vtx_task(const VTask &task, const TV *__restrict__ vdata, double *__restrict__ w,
                                         const float *__restrict__ xf, double *__restrict__ zpart, int lane) {
  switch (task.k) {
#define DFN_VTX_CASE(K) \
  case K:                                                                         \
    if constexpr (K <= KMAX) vtx_body<K, TV, ACC64, GROUPS, NT>(task, vdata, w, xf, zpart, w, lane); \
    break;
    DFN_VTX_CASE(1) DFN_VTX_CASE(2) DFN_VTX_CASE(3) DFN_VTX_CASE(4) DFN_VTX_CASE(5) DFN_VTX_CASE(6)
    DFN_VTX_CASE(7) DFN_VTX_CASE(8) DFN_VTX_CASE(9) DFN_VTX_CASE(10) DFN_VTX_CASE(11) DFN_VTX_CASE(12)
    DFN_VTX_CASE(13) DFN_VTX_CASE(14) DFN_VTX_CASE(15) DFN_VTX_CASE(16) DFN_VTX_CASE(17) DFN_VTX_CASE(18)
    DFN_VTX_CASE(19) DFN_VTX_CASE(20) DFN_VTX_CASE(21) DFN_VTX_CASE(22) DFN_VTX_CASE(23) DFN_VTX_CASE(24)
    DFN_VTX_CASE(25) DFN_VTX_CASE(26) DFN_VTX_CASE(27) DFN_VTX_CASE(28) DFN_VTX_CASE(29) DFN_VTX_CASE(30)
    DFN_VTX_CASE(31) DFN_VTX_CASE(32)
#undef DFN_VTX_CASE
    default: break;
  }
}

// Two single-strip tasks of rank K on one wave: lanes 0-31 take task a, lanes 32-63
// task b (each lane its task's fields), every lane the same kVec rows of its task as
// lane (lane & 31) of a wave of its own would; the halves are reduced apart.  A
// task's sums are those of vtx_body: there the lanes from 32 on hold zeros for a
// task of at most 32 row vectors, and the reduction's first step adds them.
template <int K, typename TV, bool ACC64, bool GROUPS, bool NT = false>
__device__ __forceinline__ void vtx_pair_body(const VTask &ta, const VTask &tb, const TV *__restrict__ vdata,
                                              const double *__restrict__ w, const float *__restrict__ xf,
                                              double *__restrict__ zpart, double *__restrict__ wz, int lane) {
  constexpr int kVec = 16 / (int)sizeof(TV);
  typedef TV tvec __attribute__((ext_vector_type(kVec)));
  typedef typename std::conditional<ACC64, double, TV>::type TA;
  TA acc[K];
#pragma unroll
  for (int j = 0; j < K; j++) acc[j] = (TA)0;
  const bool hb = lane >= 32;
  const int len = hb ? tb.len : ta.len;
  const int i = (lane & 31) * kVec;
  if (i < len) {
    const int hp = (len + kVec - 1) / kVec * kVec;  // the single strip's column stride
    const int xo = hb ? tb.xcol + tb.row0 : ta.xcol + ta.row0;
    const TV *Vs = vdata + (hb ? tb.voff : ta.voff) + i;
    TA xi[kVec];
    if constexpr (sizeof(TV) == 4 && !ACC64) {
      float4 x4;
      __builtin_memcpy(&x4, xf + xo + i, sizeof(x4));
      xi[0] = x4.x; xi[1] = x4.y; xi[2] = x4.z; xi[3] = x4.w;
    } else {
      double xv[kVec];
      __builtin_memcpy(xv, w + xo + i, sizeof(xv));
#pragma unroll
      for (int u = 0; u < kVec; u++) xi[u] = (TA)xv[u];
    }
    constexpr int G = (GROUPS && ACC64 && sizeof(TV) == 4 && K > 10) ? 10 : K;
#pragma unroll
    for (int j0 = 0; j0 < K; j0 += G) {
      tvec v[G];
#pragma unroll
      for (int j = 0; j < G; j++)
        if (j0 + j < K) v[j] = load_v<NT>(reinterpret_cast<const tvec *>(Vs + (j0 + j) * hp));
#pragma unroll
      for (int j = 0; j < G; j++)
        if (j0 + j < K)
#pragma unroll
          for (int u = 0; u < kVec; u++) acc[j0 + j] += (TA)v[j][u] * xi[u];
      if constexpr (G < K) __builtin_amdgcn_sched_barrier(0);
    }
  }
  int col;
  const TA mine = wave_reduce_scatter<K, true>(acc, lane, col);
  if (col >= 0) {
    const double v = (double)mine * (hb ? tb.scale : ta.scale);
    const int wout = hb ? tb.wout : ta.wout;
    if (wout >= 0) wz[wout + col] = v;
    else zpart[(hb ? tb.zp : ta.zp) + col] = v;
  }
}

template <typename TV, bool ACC64, int KMAX, bool GROUPS, bool NT = false>
__device__ __forceinline__ void vtx_pair_task(const VTask &ta, const VTask &tb, const TV *__restrict__ vdata,
                                              double *__restrict__ w, const float *__restrict__ xf,
                                              double *__restrict__ zpart, int lane) {
  switch (ta.k) {
#define DFN_VTX_CASE(K) \
  case K:                                                                                       \
    if constexpr (K <= KMAX) vtx_pair_body<K, TV, ACC64, GROUPS, NT>(ta, tb, vdata, w, xf, zpart, w, lane); \
    break;
    DFN_VTX_CASE(1) DFN_VTX_CASE(2) DFN_VTX_CASE(3) DFN_VTX_CASE(4) DFN_VTX_CASE(5) DFN_VTX_CASE(6)
    DFN_VTX_CASE(7) DFN_VTX_CASE(8) DFN_VTX_CASE(9) DFN_VTX_CASE(10) DFN_VTX_CASE(11) DFN_VTX_CASE(12)
    DFN_VTX_CASE(13) DFN_VTX_CASE(14) DFN_VTX_CASE(15) DFN_VTX_CASE(16) DFN_VTX_CASE(17) DFN_VTX_CASE(18)
    DFN_VTX_CASE(19) DFN_VTX_CASE(20) DFN_VTX_CASE(21) DFN_VTX_CASE(22) DFN_VTX_CASE(23) DFN_VTX_CASE(24)
    DFN_VTX_CASE(25) DFN_VTX_CASE(26) DFN_VTX_CASE(27) DFN_VTX_CASE(28) DFN_VTX_CASE(29) DFN_VTX_CASE(30)
    DFN_VTX_CASE(31) DFN_VTX_CASE(32)
#undef DFN_VTX_CASE
    default: break;
  }
}

// Phase 1 by wave items: one task, or two short tasks of the same rank and network
// (VItem), per wave.
template <typename TV, bool ACC64, int KMAX, bool NT>
__device__ __forceinline__ void vtx_items_wave(const VTask *__restrict__ tasks, const VItem *__restrict__ items,
                                               int nitems, const TV *__restrict__ vdata, double *__restrict__ w,
                                               const float *__restrict__ xf, double *__restrict__ zpart,
                                               const int *__restrict__ active) {
  const int t = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (t >= nitems) return;
  const VItem it = items[t];
  const VTask ta = tasks[it.a];
  if (active && !active[ta.net]) return;  // the items' network sits out this evaluation
  if (it.b < 0) {
    vtx_task<TV, ACC64, KMAX, true, NT>(ta, vdata, w, xf, zpart, lane);
  } else {
    const VTask tb = tasks[it.b];
    vtx_pair_task<TV, ACC64, KMAX, true, NT>(ta, tb, vdata, w, xf, zpart, lane);
  }
}

// one task per wave, blockDim.x / 64 waves per workgroup; the task index is
// uniform across the wave: scalar loads of the task and a scalar branch on its rank
template <typename TV, bool ACC64, int KMAX, bool NT>
__device__ __forceinline__ void vtx_tasks_wave(const VTask *__restrict__ tasks, int ntasks,
                                               const TV *__restrict__ vdata, double *__restrict__ w,
                                               const float *__restrict__ xf, double *__restrict__ zpart,
                                               const int *__restrict__ active) {
  const int t = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (t >= ntasks) return;
  const VTask task = tasks[t];
  if (active && !active[task.net]) return;  // the block's network sits out this evaluation
  vtx_task<TV, ACC64, KMAX, true, NT>(task, vdata, w, xf, zpart, lane);
}

template <typename TV, bool ACC64, int KMAX>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(KMAX <= 20 ? 5 : 2)))
lowrank_vtx_items_kernel(const VTask *__restrict__ tasks, const VItem *__restrict__ items, int nitems,
                         const TV *__restrict__ vdata, double *__restrict__ w, const float *__restrict__ xf,
                         double *__restrict__ zpart, const int *__restrict__ active) {
  vtx_items_wave<TV, ACC64, KMAX, false>(tasks, items, nitems, vdata, w, xf, zpart, active);
}

template <typename TV, bool ACC64, int KMAX>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(KMAX <= 20 ? 5 : 2)))
lowrank_vtx_kernel(const VTask *__restrict__ tasks, int ntasks, const TV *__restrict__ vdata, double *__restrict__ w,
                   const float *__restrict__ xf, double *__restrict__ zpart, const int *__restrict__ active) {
  vtx_tasks_wave<TV, ACC64, KMAX, false>(tasks, ntasks, vdata, w, xf, zpart, active);
}

// the same two kernels for FP64 factors read with non-temporal loads
template <int KMAX>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(KMAX <= 20 ? 5 : 2)))
lowrank_vtx_items_nt_kernel(const VTask *__restrict__ tasks, const VItem *__restrict__ items, int nitems,
                            const double *__restrict__ vdata, double *__restrict__ w, const float *__restrict__ xf,
                            double *__restrict__ zpart, const int *__restrict__ active) {
  vtx_items_wave<double, true, KMAX, true>(tasks, items, nitems, vdata, w, xf, zpart, active);
}

template <int KMAX>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(KMAX <= 20 ? 5 : 2)))
lowrank_vtx_nt_kernel(const VTask *__restrict__ tasks, int ntasks, const double *__restrict__ vdata,
                      double *__restrict__ w, const float *__restrict__ xf, double *__restrict__ zpart,
                      const int *__restrict__ active) {
  vtx_tasks_wave<double, true, KMAX, true>(tasks, ntasks, vdata, w, xf, zpart, active);
}

// z of each multi-chunk low-rank block = sum of its chunk partials, in chunk
// order; one thread per (block, column).  The partials are loaded eight at a
// time before they are added, so their latencies overlap.
__global__ void __launch_bounds__(256) zreduce_kernel(const ZRed *__restrict__ zr, int nblocks,
                                                      const double *__restrict__ zpart, double *__restrict__ w) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = g >> 5, j = g & 31;
  if (b >= nblocks) return;
  const ZRed r = zr[b];
  if (j >= r.k) return;
  const double *src = zpart + r.zp_first + j;
  double s = 0.0;
  for (int t0 = 0; t0 < r.ntask; t0 += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = t0 + u < r.ntask ? src[(int64_t)(t0 + u) * r.k] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (t0 + u < r.ntask) s += v[u];
  }
  w[r.zout + j] = s;
}

typedef double dv2 __attribute__((ext_vector_type(2)));

// Phase 2: one workgroup per panel task.  w entries of the panel's columns are
// gathered into LDS once; each wave then streams R whole rows at a time
// (row-major panel, rows padded with zero entries to a multiple of 128 bytes, so
// every row starts on a cache line and every lane issues one 16-byte load per row:
// 1 KiB per wave instruction).
template <int R>
__device__ __forceinline__ void panel_f64_body(const PTask &task, const double *__restrict__ pdata,
                                               const int *__restrict__ widx, const double *__restrict__ w,
                                               double *__restrict__ ypart, double2 *wl2) {
  double *wl = reinterpret_cast<double *>(wl2);
  for (int j = threadIdx.x; j < task.L; j += 256) wl[j] = w[widx[task.widx + j]];
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int L2 = task.L >> 1;
  const dv2 *P = reinterpret_cast<const dv2 *>(pdata + task.poff);
  // R rows per wave step (each row's sum is the same whatever R is)
  int r = wave * R;
  for (; r + R - 1 < task.m; r += 4 * R) {
    const dv2 *row0 = P + (int64_t)r * L2;
    double a[R];
#pragma unroll
    for (int q = 0; q < R; q++) a[q] = 0.0;
    // the panel is streamed once per application: non-temporal loads keep it
    // from displacing the gathered vectors in the caches
#pragma unroll 4
    for (int j = lane; j < L2; j += 64) {
      const double2 wv = wl2[j];
      dv2 v[R];
#pragma unroll
      for (int q = 0; q < R; q++) v[q] = __builtin_nontemporal_load(row0 + (int64_t)q * L2 + j);
#pragma unroll
      for (int q = 0; q < R; q++) a[q] = fma(v[q].y, wv.y, fma(v[q].x, wv.x, a[q]));
    }
    store_row_sums<R>(a, lane, ypart + task.yp + r, R);
  }
  for (; r < task.m; r++) {  // fewer than R rows left for this wave
    const dv2 *row0 = P + (int64_t)r * L2;
    double a0[1] = {0.0};
    for (int j = lane; j < L2; j += 64) {
      const double2 wv = wl2[j];
      const dv2 v0 = __builtin_nontemporal_load(row0 + j);
      a0[0] = fma(v0.y, wv.y, fma(v0.x, wv.x, a0[0]));  // as the R-row loop
    }
    store_row_sums<1>(a0, lane, ypart + task.yp + r, 1);
  }
}

template <int R>
__global__ void __launch_bounds__(256) panel_kernel_f64(const PTask *__restrict__ tasks, const double *__restrict__ pdata,
                                                        const int *__restrict__ widx, const double *__restrict__ w,
                                                        double *__restrict__ ypart, const int *__restrict__ active) {
  __shared__ double2 wl2[kPanelMaxCols / 2];
  const PTask task = tasks[blockIdx.x];
  if (active && !active[task.net]) return;  // the panel's network sits out this evaluation
  panel_f64_body<R>(task, pdata, widx, w, ypart, wl2);
}

// Phase 1 over FP32 low-rank factors (FP64 products and sums) and phase 2 of the
// FP64 near-field panels of a mixed-precision layout, in one launch: the
// near-field panels read only x_h, so they stream beside the low-rank tasks
// instead of after them.  Workgroup b is a panel workgroup when
// floor((b + 1) np / nb) > floor(b np / nb) (nb workgroups in all), which spreads
// the panels evenly over the grid; every other workgroup runs four phase-1 tasks.
// Each sum is computed as the two separate kernels compute it.  Four waves per SIMD
// (128 VGPRs) for either kind of workgroup.
template <int R, int KMAX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) vtx_near_kernel(const VTask *__restrict__ vt, int nvt,
                                                       const float *__restrict__ vdata, const PTask *__restrict__ pt,
                                                       int npt, const double *__restrict__ pdata,
                                                       const int *__restrict__ widx, double *__restrict__ w,
                                                       double *__restrict__ zpart, double *__restrict__ ypart,
                                                       const int *__restrict__ active) {
  __shared__ double2 wl2[kPanelMaxCols / 2];
  const int64_t nb = (int64_t)((nvt + 3) >> 2) + npt, b = blockIdx.x;
  const int p0 = (int)(b * npt / nb), p1 = (int)((b + 1) * npt / nb);
  if (p1 > p0) {
    const PTask task = pt[p0];
    if (active && !active[task.net]) return;
    panel_f64_body<R>(task, pdata, widx, w, ypart, wl2);
    return;
  }
  const int t = __builtin_amdgcn_readfirstlane((int)(b - p1) * 4 + (int)(threadIdx.x >> 6));
  if (t >= nvt) return;
  const VTask task = vt[t];
  if (active && !active[task.net]) return;
  // held to 128 VGPRs as a whole; the column groups would only constrain its schedule
  vtx_task<float, true, KMAX, false>(task, vdata, w, nullptr, zpart, threadIdx.x & 63);
}

// FP32 storage: entries are block / 2^e per block, the column scales 2^e fold into
// the gathered input; FP32 accumulation, with a compensated (Kahan) sum over the
// near-field columns j < nd of each row.  Lane partials are combined in FP64.
// (contraction off: the compensation needs the rounded difference and sum, and every
// launch shape of the panel must compute the same floats)
__device__ __forceinline__ void kahan_add(float &s, float &c, float v) {
#pragma clang fp contract(off)
  const float y = v - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

typedef float fv4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void acc4(fv4 v, float4 wv, int j, int nd, float &s, float &c, float &a) {
#pragma clang fp contract(off)
  if (j + 3 < nd) {
    kahan_add(s, c, v.x * wv.x);
    kahan_add(s, c, v.y * wv.y);
    kahan_add(s, c, v.z * wv.z);
    kahan_add(s, c, v.w * wv.w);
  } else if (j >= nd) {
    a = fmaf(v.w, wv.w, fmaf(v.z, wv.z, fmaf(v.y, wv.y, fmaf(v.x, wv.x, a))));
  } else {
    const float p[4] = {v.x * wv.x, v.y * wv.y, v.z * wv.z, v.w * wv.w};
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if (j + q < nd) kahan_add(s, c, p[q]);
      else a += p[q];
    }
  }
}

template <int R>
__global__ void __launch_bounds__(256) panel_kernel_f32(const PTask *__restrict__ tasks, const float *__restrict__ pdata,
                                                        const int *__restrict__ widx, const float *__restrict__ cscale,
                                                        const double *__restrict__ w, double *__restrict__ ypart,
                                                        const int *__restrict__ active) {
  __shared__ float4 wl4[kPanelMaxCols / 4];
  float *wl = reinterpret_cast<float *>(wl4);
  const PTask task = tasks[blockIdx.x];
  if (active && !active[task.net]) return;  // the panel's network sits out this evaluation
  for (int j = threadIdx.x; j < task.L; j += 256)
    wl[j] = (float)(w[widx[task.widx + j]] * (double)cscale[task.widx + j]);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int L4 = task.L >> 2, nd = task.nd;
  const fv4 *P = reinterpret_cast<const fv4 *>(pdata + task.poff);
  // R rows per wave step (rows past the panel re-read its first row of the step and
  // are not stored; each row's sum is the same whatever R is)
  for (int r = wave * R; r < task.m; r += 4 * R) {
    const fv4 *row[R];
    float s[R], c[R], a[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
      row[q] = P + (int64_t)(r + q < task.m ? r + q : r) * L4;
      s[q] = 0.0f; c[q] = 0.0f; a[q] = 0.0f;
    }
#pragma unroll 2
    for (int j = lane; j < L4; j += 64) {
      const float4 wv = wl4[j];
      fv4 v[R];
#pragma unroll
      for (int q = 0; q < R; q++) v[q] = __builtin_nontemporal_load(row[q] + j);
#pragma unroll
      for (int q = 0; q < R; q++) acc4(v[q], wv, 4 * j, nd, s[q], c[q], a[q]);
    }
    double t[R];
#pragma unroll
    for (int q = 0; q < R; q++) t[q] = ((double)s[q] - (double)c[q]) + (double)a[q];
    store_row_sums<R>(t, lane, ypart + task.yp + r, task.m - r);
  }
}

// Phase 2 over FP32 storage with FP64 arithmetic: the gathered entries stay FP64
// (in 16 KB of LDS), every stored entry is widened exactly and multiplied and
// summed in FP64, so the only rounding left beside FP64 is that of the storage.
template <int R>
__global__ void __launch_bounds__(256) panel_kernel_f32w(const PTask *__restrict__ tasks,
                                                         const float *__restrict__ pdata, const int *__restrict__ widx,
                                                         const float *__restrict__ cscale, const double *__restrict__ w,
                                                         double *__restrict__ ypart, const int *__restrict__ active) {
  __shared__ double2 wl2[kPanelMaxCols / 2];
  double *wl = reinterpret_cast<double *>(wl2);
  const PTask task = tasks[blockIdx.x];
  if (active && !active[task.net]) return;  // the panel's network sits out this evaluation
  for (int j = threadIdx.x; j < task.L; j += 256) wl[j] = w[widx[task.widx + j]] * (double)cscale[task.widx + j];
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int L4 = task.L >> 2;
  const fv4 *P = reinterpret_cast<const fv4 *>(pdata + task.poff);
  for (int r = wave * R; r < task.m; r += 4 * R) {
    const fv4 *row[R];
    double a[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
      row[q] = P + (int64_t)(r + q < task.m ? r + q : r) * L4;
      a[q] = 0.0;
    }
#pragma unroll 2
    for (int j = lane; j < L4; j += 64) {
      const double2 w01 = wl2[2 * j], w23 = wl2[2 * j + 1];
      fv4 v[R];
#pragma unroll
      for (int q = 0; q < R; q++) v[q] = __builtin_nontemporal_load(row[q] + j);
#pragma unroll
      for (int q = 0; q < R; q++)
        a[q] = fma((double)v[q].w, w23.y, fma((double)v[q].z, w23.x, fma((double)v[q].y, w01.y,
                   fma((double)v[q].x, w01.x, a[q]))));
    }
    store_row_sums<R>(a, lane, ypart + task.yp + r, task.m - r);
  }
}

// Phase 2 over near-field entries stored as an FP32 value and an 8-bit residual:
// entry = hi + q * s, with hi = (float)entry, q the rounding of (entry - hi) / s to an
// integer in [-127, 127] and s = (the task's largest |entry - hi|) / 127 (cscale at the
// task's first column).  Two streams: the FP32 values, rows of L entries (256-byte
// multiples), and the residual bytes, rows of L bytes (so rows 2i and 2i + 1, which
// the same wave step reads, fill whole 128-byte lines).  The hi products and sums are
// FP64; the residual terms, each below 2^-24 of its entry, are summed per lane in
// FP32 against (float)w, scaled by s and added to the FP64 sum before the wave's
// reduction.  The largest entries of a task keep about 31 significant bits, every
// entry at least FP32's 24, in 5 bytes.  Tasks of at most kSplitMaxCols columns keep
// w (FP64 and FP32) in 12 KB of LDS.
template <int R>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(R == 2 ? 6 : 4)))
panel_kernel_split(const PTask *__restrict__ tasks, const float *__restrict__ hi, const signed char *__restrict__ lo,
                   const int *__restrict__ widx, const float *__restrict__ cscale, const double *__restrict__ w,
                   double *__restrict__ ypart, const int *__restrict__ active) {
  __shared__ double2 wl2[kSplitMaxCols / 2];
  __shared__ fv4 wf4[kSplitMaxCols / 4];
  double *wl = reinterpret_cast<double *>(wl2);
  float *wf = reinterpret_cast<float *>(wf4);
  const PTask task = tasks[blockIdx.x];
  if (active && !active[task.net]) return;  // the panel's network sits out this evaluation
  for (int j = threadIdx.x; j < task.L; j += 256) {
    const double v = w[widx[task.widx + j]];
    wl[j] = v;
    wf[j] = (float)v;
  }
  const double s = (double)cscale[task.widx];
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int L4 = task.L >> 2;
  const fv4 *H = reinterpret_cast<const fv4 *>(hi + task.poff);
  const unsigned *Q = reinterpret_cast<const unsigned *>(lo + task.poff);
  for (int r = wave * R; r < task.m; r += 4 * R) {
    const fv4 *hrow[R];
    const unsigned *qrow[R];
    double a[R];
    float b[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
      const int rr = r + q < task.m ? r + q : r;  // rows past the panel re-read its row r and are not stored
      hrow[q] = H + (int64_t)rr * L4;
      qrow[q] = Q + (int64_t)rr * L4;
      a[q] = 0.0;
      b[q] = 0.0f;
    }
#pragma unroll 1
    for (int j = lane; j < L4; j += 64) {
      const double2 w01 = wl2[2 * j], w23 = wl2[2 * j + 1];
      const fv4 f = wf4[j];
      fv4 h[R];
      unsigned l[R];
#pragma unroll
      for (int q = 0; q < R; q++) {
        h[q] = __builtin_nontemporal_load(hrow[q] + j);
        l[q] = __builtin_nontemporal_load(qrow[q] + j);
      }
#pragma unroll
      for (int q = 0; q < R; q++) {
        a[q] = fma((double)h[q].w, w23.y, fma((double)h[q].z, w23.x, fma((double)h[q].y, w01.y,
                   fma((double)h[q].x, w01.x, a[q]))));
        const float q0 = (float)(signed char)l[q], q1 = (float)(signed char)(l[q] >> 8);
        const float q2 = (float)(signed char)(l[q] >> 16), q3 = (float)(signed char)(l[q] >> 24);
        b[q] = fmaf(q3, f.w, fmaf(q2, f.z, fmaf(q1, f.y, fmaf(q0, f.x, b[q]))));
      }
    }
    double t[R];
#pragma unroll
    for (int q = 0; q < R; q++) t[q] = fma((double)b[q], s, a[q]);
    store_row_sums<R>(t, lane, ypart + task.yp + r, task.m - r);
  }
}

// y(perm(i)) = scale * sum of the partials of H row i (one thread per row and operator).
__global__ void __launch_bounds__(256) finalize_kernel(const int *__restrict__ fin_first, const int *__restrict__ fin_idx,
                                                       const int *__restrict__ perm, const double *__restrict__ ypart,
                                                       int n, int nops, double scale, double *__restrict__ yT,
                                                       double *__restrict__ yN) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n * nops) return;
  const int op = g / n, i = g - op * n;
  const int q0 = fin_first[g], q1 = fin_first[g + 1];
  double s = 0.0;
  for (int q = q0; q < q1; q += 4) {  // four partials in flight, added in order
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) v[u] = q + u < q1 ? ypart[fin_idx[q + u]] : 0.0;
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (q + u < q1) s += v[u];
  }
  double *y = op == 0 ? yT : yN;
  y[perm[i]] = scale * s;
}

// ---------------------------------------------------------------------------
// Multi-RHS stress transfer: the H-matrix applied to R slip-rate vectors at once,
// both phases as dense products on the matrix cores (v_mfma_f64_16x16x4f64) over
// tiles of 16 vectors.  Every vector set is row-major with a stride R that is a
// multiple of 64: the gathered input and the z values W[q R + r], the phase-1
// partials zpart[q R + r], the panel row partials ypart[q R + r].
// Operand layout of the FP64 16x16x4 product: lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16]; accumulator element e of lane l is C[l / 16 + 4 e][l % 16].
// ---------------------------------------------------------------------------
typedef double d4v __attribute__((ext_vector_type(4)));

// W[i R + r] = X[r n + perm[i]] (r < nrhs; 0 for the padding vectors), in 32 x 32 tiles
__global__ void __launch_bounds__(256) mrhs_gather_kernel(const double *__restrict__ X, const int *__restrict__ perm,
                                                          int n, int nrhs, int R, double *__restrict__ W) {
  __shared__ double t[32][33];
  const int i0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int q = ty; q < 32; q += 8) {
    const int i = i0 + tx, r = r0 + q;
    t[q][tx] = (i < n && r < nrhs) ? X[(int64_t)r * n + perm[i]] : 0.0;
  }
  __syncthreads();
  for (int q = ty; q < 32; q += 8) {
    const int i = i0 + q, r = r0 + tx;
    if (i < n) W[(int64_t)i * R + r] = t[tx][q];
  }
}

// Phase 1: z_j[r] = scale * sum_i V[i][j] X[xcol + row0 + i][r] for j < k, one wave
// per (task, group of 64 vectors): products of the task's V^T (k x len, in K
// tiles of 16) with its X rows (len x 64, four tiles of 16), 4 rows per product.
// The loads of 16 rows are issued together before their products.
__global__ void __launch_bounds__(256) mrhs_vtx_kernel(const VTask *__restrict__ tasks, int ntasks,
                                                       const double *__restrict__ vdata, double *__restrict__ W,
                                                       double *__restrict__ zpart, int R, int groups) {
  const int wv = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63, am = lane & 15, ak = lane >> 4;
  const int t = wv / groups, g = wv - t * groups;
  if (t >= ntasks) return;
  const VTask task = tasks[t];
  const int K = task.k, len = task.len;
  const bool two = K > 16;
  const double *V = vdata + task.voff;
  const double *X = W + (int64_t)(task.xcol + task.row0) * R + g * 64 + am;
  d4v acc[2][4];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = d4v{0.0, 0.0, 0.0, 0.0};
  for (int i0 = 0; i0 < len; i0 += 16) {
    double a[4][2], x[4][4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int i = i0 + 4 * u + ak;
      const bool in = i < len;
      // V of a task is stored strip by strip (128 rows each for FP64), column after column
      const int s0 = (i / 128) * 128, hp = (min(128, len - s0) + 1) & ~1;
      const double *Vi = V + (int64_t)s0 * K + (i - s0);
      a[u][0] = (in && am < K) ? Vi[(int64_t)am * hp] : 0.0;
      a[u][1] = (in && two && 16 + am < K) ? Vi[(int64_t)(16 + am) * hp] : 0.0;
#pragma unroll
      for (int tile = 0; tile < 4; tile++) x[u][tile] = in ? X[(int64_t)i * R + tile * 16] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int tile = 0; tile < 4; tile++) {
        acc[0][tile] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][0], x[u][tile], acc[0][tile], 0, 0, 0);
        if (two) acc[1][tile] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][1], x[u][tile], acc[1][tile], 0, 0, 0);
      }
  }
#pragma unroll
  for (int kt = 0; kt < 2; kt++)
#pragma unroll
    for (int tile = 0; tile < 4; tile++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int j = kt * 16 + ak + 4 * e;
        if (j < K) {
          double *dst = task.wout >= 0 ? W + (int64_t)(task.wout + j) * R : zpart + (int64_t)(task.zp + j) * R;
          dst[g * 64 + tile * 16 + am] = acc[kt][tile][e] * task.scale;
        }
      }
}

// z of each multi-chunk block: the sum of its chunk partials, in chunk order
__global__ void __launch_bounds__(256) mrhs_zreduce_kernel(const ZRed *__restrict__ zr, int nblocks,
                                                           const double *__restrict__ zpart, double *__restrict__ W,
                                                           int R) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = 32 * (int64_t)R;
  const int b = (int)(g / per), j = (int)((g % per) / R), r = (int)(g % R);
  if (b >= nblocks) return;
  const ZRed z = zr[b];
  if (j >= z.k) return;
  double s = 0.0;
  for (int t = 0; t < z.ntask; t++) s += zpart[(int64_t)(z.zp_first + t * z.k + j) * R + r];
  W[(int64_t)(z.zout + j) * R + r] = s;
}

// Phase 2: ypart[yp + row][r] = sum_c P[row][c] W[widx[c]][r]; one workgroup per
// (panel task, group of 64 vectors), wave w on vectors 16 w..16 w + 15 of the group.
// The panel is staged through LDS 64 rows x 32 columns at a time (double-buffered,
// one barrier per chunk) and the loads of the next chunk (its panel entries, its
// gathered vector entries) are in flight while the current chunk's products run;
// the column indices run one chunk further ahead.  Row tiles of 16 past the
// panel's last row are skipped.
__global__ void __launch_bounds__(256) mrhs_panel_kernel(const PTask *__restrict__ tasks,
                                                         const double *__restrict__ pdata,
                                                         const int *__restrict__ widx, const double *__restrict__ W,
                                                         double *__restrict__ ypart, int R, int groups) {
  __shared__ double As[2][64][33];
  const int t = blockIdx.x / groups, g = blockIdx.x - t * groups;
  const PTask task = tasks[t];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, am = lane & 15, ak = lane >> 4;
  const int rhs = g * 64 + wave * 16 + am;
  const double *P = pdata + task.poff;
  const int *wi = widx + task.widx;
  const int m = task.m, L = task.L;
  // (a panel with no columns still writes its zero rows)
  const int nch = max(1, (L + 31) >> 5), nrb = (m + 63) >> 6, nsteps = nch * nrb;
  // step q: row block q / nch, columns 32 (q % nch) ..
  double pa[8], pb[8], cb[8];
  int pw[8];
  auto load_a = [&](int q) {
    const int row0 = (q / nch) * 64, c0 = (q % nch) * 32;
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int r = (threadIdx.x >> 5) + 8 * u, c = threadIdx.x & 31;
      pa[u] = (row0 + r < m && c0 + c < L) ? P[(int64_t)(row0 + r) * L + c0 + c] : 0.0;
    }
  };
  auto load_w = [&](int q) {
    const int c0 = (q % nch) * 32;
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int c = c0 + 4 * u + ak;
      pw[u] = c < L ? wi[c] : -1;
    }
  };
  auto load_b = [&]() {
#pragma unroll
    for (int u = 0; u < 8; u++) pb[u] = pw[u] >= 0 ? W[(int64_t)pw[u] * R + rhs] : 0.0;
  };
  load_a(0);
  load_w(0);
  load_b();
  if (nsteps > 1) load_w(1);
  d4v acc[4];
  for (int q = 0; q < nsteps; q++) {
    const int buf = q & 1, rb = q / nch, ch = q - rb * nch;
#pragma unroll
    for (int u = 0; u < 8; u++) As[buf][(threadIdx.x >> 5) + 8 * u][threadIdx.x & 31] = pa[u];
#pragma unroll
    for (int u = 0; u < 8; u++) cb[u] = pb[u];
    __syncthreads();
    if (q + 1 < nsteps) {
      load_a(q + 1);
      load_b();
      if (q + 2 < nsteps) load_w(q + 2);
    }
    if (ch == 0)
#pragma unroll
      for (int rt = 0; rt < 4; rt++) acc[rt] = d4v{0.0, 0.0, 0.0, 0.0};
    const int row0 = rb * 64;
#pragma unroll
    for (int rt = 0; rt < 4; rt++)
      if (row0 + rt * 16 < m)
#pragma unroll
        for (int u = 0; u < 8; u++)
          acc[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(As[buf][rt * 16 + am][4 * u + ak], cb[u], acc[rt], 0, 0, 0);
    if (ch == nch - 1)
#pragma unroll
      for (int rt = 0; rt < 4; rt++)
#pragma unroll
        for (int e = 0; e < 4; e++) {
          const int row = row0 + rt * 16 + ak + 4 * e;
          if (row < m) ypart[(int64_t)(task.yp + row) * R + rhs] = acc[rt][e];
        }
  }
}

// Yh[(o n + i) R + r] = sum of the partials of H row i of operator o, in order
__global__ void __launch_bounds__(256) mrhs_finalize_kernel(const int *__restrict__ fin_first,
                                                            const int *__restrict__ fin_idx,
                                                            const double *__restrict__ ypart, int n, int nops, int R,
                                                            double *__restrict__ Yh) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (int64_t)n * nops * R) return;
  const int64_t row = g / R;
  const int r = (int)(g - row * R);
  double s = 0.0;
  for (int q = fin_first[row]; q < fin_first[row + 1]; q++) s += ypart[(int64_t)fin_idx[q] * R + r];
  Yh[g] = s;
}

// y[r n + perm[i]] = scale * Yh[i R + r] (r < nrhs), in 32 x 32 tiles
__global__ void __launch_bounds__(256) mrhs_scatter_kernel(const double *__restrict__ Yh, const int *__restrict__ perm,
                                                           int n, int nrhs, int R, double scale,
                                                           double *__restrict__ y) {
  __shared__ double t[32][33];
  const int i0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int q = ty; q < 32; q += 8) {
    const int i = i0 + q, r = r0 + tx;
    t[q][tx] = i < n ? Yh[(int64_t)i * R + r] : 0.0;
  }
  __syncthreads();
  for (int q = ty; q < 32; q += 8) {
    const int i = i0 + tx, r = r0 + q;
    if (i < n && r < nrhs) y[(int64_t)r * n + perm[i]] = scale * t[tx][q];
  }
}

// ---------------------------------------------------------------------------
// Direct kernels (src/kernel_calculation/directModeIII.f:7-74, directModeII.f90:7-107)
// ---------------------------------------------------------------------------
// Dense direct kernels (src/kernel_calculation/directModeIII.f:7-74,
// directModeII.f90:7-107).  Both sum over the N+1 nodes rather than the N
// elements: the reference adds each element's contribution as (left node term)
// - (right node term), so the node terms are gathered once per node with the
// slip differences of the two elements sharing it ("dislocation weights").
// The N x (N+1) node plane is tiled as 256 receiver rows x kDirectCols nodes
// per workgroup, so a 4000-element fault still launches ~1000 workgroups; each
// tile writes one partial per row and direct_reduce_kernel sums the partials of
// a row in a fixed order (results do not depend on scheduling).  The weights
// are formed from the slip as the tiles load their nodes.
// ---------------------------------------------------------------------------
// a / b for b > 0 (a squared distance between an element centre and a node)
// without the IEEE divide sequence: the hardware reciprocal estimate refined by two
// Newton-Raphson steps, y <- y + y (1 - b y), then one product; within about one
// ulp of the rounded quotient
__device__ __forceinline__ double quot_pos(double a, double b) {
  double y = __builtin_amdgcn_rcp(b);
  double e = fma(-b, y, 1.0);
  y = fma(y, e, y);
  e = fma(-b, y, 1.0);
  y = fma(y, e, y);
  return a * y;
}

// Mode III weight of node l: -(s_l - s_{l-1}) with s_{-1} = s_N = 0
__device__ __forceinline__ double dislocation_weight3(const double *__restrict__ slip, int l, int n) {
  if (l == 0) return -slip[0];
  if (l == n) return slip[n - 1];
  return -(slip[l] - slip[l - 1]);
}

__global__ void __launch_bounds__(256) direct3_tile_kernel(const double *__restrict__ node,
                                                           const double *__restrict__ slip, int n,
                                                           double *__restrict__ part) {
  __shared__ double sx[kDirectCols], sy[kDirectCols], sc[kDirectCols];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j0 = blockIdx.y * kDirectCols, cnt = min(kDirectCols, n + 1 - j0);
  if (threadIdx.x < cnt) {
    const int l = j0 + threadIdx.x;
    sx[threadIdx.x] = node[2 * l];
    sy[threadIdx.x] = node[2 * l + 1];
    sc[threadIdx.x] = dislocation_weight3(slip, l, n);
  }
  __syncthreads();
  if (k >= n) return;
  const double x0 = node[2 * k], y0 = node[2 * k + 1], x1 = node[2 * k + 2], y1 = node[2 * k + 3];
  const double tx = x1 - x0, ty = y1 - y0;
  const double nt = sqrt(tx * tx + ty * ty);
  const double t1 = tx / nt, t2 = ty / nt;
  const double ex = x0 + 0.5 * (x1 - x0), ey = y0 + 0.5 * (y1 - y0);
  double s = 0.0;
  for (int q = 0; q < cnt; q++) {
    const double dx = ex - sx[q], dy = ey - sy[q];
    const double r2 = dx * dx + dy * dy;
    s += sc[q] * quot_pos(t1 * dx + t2 * dy, r2);
  }
  part[(int64_t)blockIdx.y * n + k] = s;
}

// Mode II weights of node l: w1 = s_l m1_l - s_{l-1} m1_{l-1}, w2 likewise with m2
// (m = element normal, s = slip; s_{-1} = s_N = 0)
__global__ void __launch_bounds__(256) direct2_tile_kernel(const double *__restrict__ node,
                                                           const double *__restrict__ elem,
                                                           const double *__restrict__ nrm,
                                                           const double *__restrict__ slip, int n,
                                                           double *__restrict__ partT, double *__restrict__ partN) {
  __shared__ double sx[kDirectCols], sy[kDirectCols], sw1[kDirectCols], sw2[kDirectCols];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int j0 = blockIdx.y * kDirectCols, cnt = min(kDirectCols, n + 1 - j0);
  if (threadIdx.x < cnt) {
    const int l = j0 + threadIdx.x;
    sx[threadIdx.x] = node[2 * l];
    sy[threadIdx.x] = node[2 * l + 1];
    double w1 = 0.0, w2 = 0.0;
    if (l < n) { w1 = slip[l] * nrm[2 * l]; w2 = slip[l] * nrm[2 * l + 1]; }
    if (l > 0) { w1 -= slip[l - 1] * nrm[2 * l - 2]; w2 -= slip[l - 1] * nrm[2 * l - 1]; }
    sw1[threadIdx.x] = w1;
    sw2[threadIdx.x] = w2;
  }
  __syncthreads();
  if (k >= n) return;
  const double ex = elem[2 * k], ey = elem[2 * k + 1], n1 = nrm[2 * k], n2 = nrm[2 * k + 1];
  const double c4 = 4.0 * n1 * n2, c2 = 2.0 * n1 * n2, dn = n2 * n2 - n1 * n1;
  double aT = 0.0, aN = 0.0;
  for (int q = 0; q < cnt; q++) {
    const double gx = ex - sx[q], gy = ey - sy[q];
    const double r = sqrt(gx * gx + gy * gy), ir = 1.0 / r;
    const double g1 = gx * ir, g2 = gy * ir;
    const double d21 = g2 * g2 - g1 * g1, p12 = g1 * g2;
    const double cc = c4 * p12 + dn * d21;
    const double dd = c2 * d21 - dn * 2.0 * p12;
    const double I1t = cc * g1 * ir, I2t = cc * g2 * ir;
    const double I1n = g1 * ir - g2 * ir * dd, I2n = g2 * ir + g1 * ir * dd;
    const double w1 = sw1[q], w2 = sw2[q];
    aT += I1t * w2 - I2t * w1;
    aN += I1n * w1 + I2n * w2;
  }
  partT[(int64_t)blockIdx.y * n + k] = aT;
  partN[(int64_t)blockIdx.y * n + k] = aN;
}

// out[k] = scale * sum_c part[c][k]: 64 rows per workgroup, four waves each
// summing every fourth tile, then the four sums added in wave order
__global__ void __launch_bounds__(256) direct_reduce_kernel(const double *__restrict__ part, int nchunk, int n,
                                                            double scale, double *__restrict__ out) {
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int k = blockIdx.x * 64 + lane;
  double s = 0.0;
  if (k < n) {
#pragma unroll 4
    for (int c = g; c < nchunk; c += 4) s += part[(int64_t)c * n + k];
  }
  red[g][lane] = s;
  __syncthreads();
  if (g == 0 && k < n) out[k] = (((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane]) * scale;
}

// ---------------------------------------------------------------------------
// Rate-state RHS fused with the Cash-Karp stage combinations
// (src/friction/RateStateAgingModeIII.f90:66-113, ModeII.f90:169-212,
//  src/ode_solver/ode_bs.f90:199-224)
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)

struct CK {
  static constexpr double A2 = 0.2, A3 = 0.3, A4 = 0.6, A5 = 1.0, A6 = 0.875;
  static constexpr double B21 = 0.2, B31 = 3.0 / 40.0, B32 = 9.0 / 40.0, B41 = 0.3, B42 = -0.9, B43 = 1.2;
  static constexpr double B51 = -11.0 / 54.0, B52 = 2.5, B53 = -70.0 / 27.0, B54 = 35.0 / 27.0;
  static constexpr double B61 = 1631.0 / 55296.0, B62 = 175.0 / 512.0, B63 = 575.0 / 13824.0;
  static constexpr double B64 = 44275.0 / 110592.0, B65 = 253.0 / 4096.0;
  static constexpr double C1 = 37.0 / 378.0, C3 = 250.0 / 621.0, C4 = 125.0 / 594.0, C6 = 512.0 / 1771.0;
  static constexpr double DC1 = C1 - 2825.0 / 27648.0, DC3 = C3 - 18575.0 / 48384.0;
  static constexpr double DC4 = C4 - 13525.0 / 55296.0, DC5 = -277.0 / 14336.0, DC6 = C6 - 0.25;
};

__device__ __forceinline__ void rate_state(const RhsParams &p, int k, double V, double psi, double sig, double sT,
                                           double sN, double Pt, double Pd, double &dV, double &dpsi, double &dsig) {
  dsig = p.mode3 ? p.nld[k] : sN + p.nld[k];
  const double a = p.a[k], b = p.b[k], Dc = p.Dc[k], V0 = p.V0[k];
  const double B1 = V / (2.0 * V0);
  const double C1 = exp((p.f0[k] + b * psi) / a);
  double D1 = 1.0 / (C1 * C1) + B1 * B1;
  D1 = 1.0 / sqrt(D1);
  const double th = V0 / Dc * exp(-psi) - fabs(V) / Dc;
  dV = (p.sld[k] + sT + (dsig + Pd) * a * asinh(B1 * C1) + B1 * D1 * (sig + Pt) * b * th) /
       (p.eta_rad - D1 * (sig + Pt) * a / (2.0 * V0));
  dpsi = V0 / Dc * exp(-psi) - fabs(V) / Dc;
  if (p.fake[k]) { dV = 0.0; dpsi = 0.0; dsig = 0.0; }
}

// Slip rate from the quasi-dynamic stress balance whose time derivative the
// reference integrates (src/friction/RateStateAgingModeII.f90:61-66):
// g(V) = eta V - s a asinh(V C / (2 V0)) - tau = 0, s = sigma + P < 0 (compressive),
// C = exp((f0 + b psi) / a).  g is odd in (V, tau) and increasing in V, so the root
// has the sign of tau; |V| is found by Newton steps in ln V (the friction term is
// logarithmic in V over most of its range), kept inside the bracket [lo, hi] with
// g(lo) <= 0 <= g(hi), starting from the last solution.  Without compressive
// normal stress the balance has no friction term: V = tau / eta.
__device__ double nv_solve_V(double tau, double s, double a, double C, double V0, double eta, double Vg) {
  if (tau == 0.0) return 0.0;
  const double t = fabs(tau);
  if (!(s < 0.0)) return tau / eta;
  double lo = 0.0, hi = t / eta;
  double V = fabs(Vg);
  if (!(V > lo && V < hi)) V = 0.5 * hi;
  for (int it = 0; it < 200; it++) {
    const double B = V / (2.0 * V0);
    const double g = eta * V - s * a * asinh(B * C) - t;
    if (g == 0.0) break;
    if (g > 0.0) hi = V;
    else lo = V;
    const double D = 1.0 / sqrt(1.0 / (C * C) + B * B);  // (C / 2V0) / sqrt(1 + (B C)^2) = D / 2V0
    const double gp = eta - s * a * D / (2.0 * V0);
    double Vn = V * exp(-g / (gp * V));
    if (!(Vn > lo && Vn < hi)) Vn = lo > 0.0 ? sqrt(lo * hi) : 0.5 * (lo + hi);
    const bool done = fabs(Vn - V) <= 1e-14 * Vn;
    V = Vn;
    if (done) break;
  }
  return tau < 0.0 ? -V : V;
}

// Newton-V right-hand side at element k: [dslip, dpsi, 0] and the normal stress
// of this evaluation (for the tensile check).  The aging law is the reference's
// (RateStateAgingModeII.f90:69-70).
__device__ __forceinline__ void rate_state_nv(const RhsParams &p, const NvArgs &nv, int k, double psi, double sT,
                                              double sN, double Pt, double &dd, double &dpsi, double &dz,
                                              double &sig) {
  const double a = p.a[k], b = p.b[k], Dc = p.Dc[k], V0 = p.V0[k];
  sig = nv.sig_ref[k] + p.nld[k] * nv.t + (p.mode3 ? 0.0 : sN);
  const double tau = nv.tau_ref[k] + p.sld[k] * nv.t + sT;
  const double C = exp((p.f0[k] + b * psi) / a);
  const double V = nv_solve_V(tau, sig + Pt, a, C, V0, p.eta_rad, nv.Vg[k]);
  dd = V;
  dpsi = V0 / Dc * exp(-psi) - fabs(V) / Dc;
  dz = 0.0;
  if (p.fake[k]) { dd = 0.0; dpsi = 0.0; sig = -1.0; }
  else nv.Vg[k] = V;
}

__global__ void __launch_bounds__(256) nv_enter_kernel(RhsParams p, double *Y, const double *P, int n,
                                                       double *tau_ref, double *sig_ref, double *Vg) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double V = Y[k], psi = Y[n + k], sig = Y[2 * n + k];
  const double a = p.a[k], V0 = p.V0[k];
  const double C = exp((p.f0[k] + p.b[k] * psi) / a);
  tau_ref[k] = p.eta_rad * V - (sig + P[k]) * a * asinh(V / (2.0 * V0) * C);
  sig_ref[k] = sig;
  Vg[k] = V;
  Y[k] = 0.0;
  Y[2 * n + k] = 0.0;
}

__global__ void __launch_bounds__(256) nv_state_kernel(RhsParams p, NvArgs nv, const double *Y, const double *P,
                                                       const double *sT, const double *sN, int n, double *out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double psi = Y[n + k];
  const double a = p.a[k], V0 = p.V0[k];
  const double sig = nv.sig_ref[k] + p.nld[k] * nv.t + (p.mode3 ? 0.0 : sN[k]);
  const double tau = nv.tau_ref[k] + p.sld[k] * nv.t + sT[k];
  const double C = exp((p.f0[k] + p.b[k] * psi) / a);
  out[k] = p.fake[k] ? nv.Vg[k] : nv_solve_V(tau, sig + P[k], a, C, V0, p.eta_rad, nv.Vg[k]);
  out[n + k] = psi;
  out[2 * n + k] = sig;
}

void launch_nv_enter(const RhsParams &p, double *Y, const double *P, int n, double *tau_ref, double *sig_ref,
                     double *Vg, hipStream_t s) {
  hipLaunchKernelGGL(nv_enter_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, Y, P, n, tau_ref, sig_ref, Vg);
}

void launch_nv_state(const RhsParams &p, const NvArgs &nv, const double *Y, const double *P, const double *sT,
                     const double *sN, int n, double *out, hipStream_t s) {
  hipLaunchKernelGGL(nv_state_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, nv, Y, P, sT, sN, n, out);
}

// Stage s of rkck at element k: k_s = f(ytemp_s); then the next stage input (or,
// at s = 6, the 5th-order solution, and the scaled embedded error into emax).  At
// s = 1 it evaluates k1 at x = t_try and forms yscal (src/solver.f90:101-102).
template <int S>
__device__ __forceinline__ void stage_body(const StageArgs &A, int k, double h, double t_try, double tol, double eta,
                                           double &emax, int &tensile) {
  const int n = A.n;
  const double *yt = A.ytemp;
  double d[3];
  RhsParams p = A.p;
  p.eta_rad = eta;
  double sig_nv = -1.0;
  if (A.nv.on)
    rate_state_nv(p, A.nv, k, yt[n + k], A.sT[k], A.sN ? A.sN[k] : 0.0, A.Pt[k], d[0], d[1], d[2], sig_nv);
  else
    rate_state(p, k, yt[k], yt[n + k], yt[2 * n + k], A.sT[k], A.sN ? A.sN[k] : 0.0, A.Pt[k], A.Pd[k], d[0], d[1],
               d[2]);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const int i = c * n + k;
    const double y = A.y[i];
    if (S == 0) {
      A.k[0][i] = d[c];
    } else if (S == 1) {
      A.k[0][i] = d[c];
      A.yscal[i] = fabs(y) + fabs(d[c] * t_try) + tol;
      A.ynext[i] = y + CK::B21 * h * d[c];
    } else if (S == 2) {
      A.k[1][i] = d[c];
      A.ynext[i] = y + h * (CK::B31 * A.k[0][i] + CK::B32 * d[c]);
    } else if (S == 3) {
      A.k[2][i] = d[c];
      A.ynext[i] = y + h * (CK::B41 * A.k[0][i] + CK::B42 * A.k[1][i] + CK::B43 * d[c]);
    } else if (S == 4) {
      A.k[3][i] = d[c];
      A.ynext[i] = y + h * (CK::B51 * A.k[0][i] + CK::B52 * A.k[1][i] + CK::B53 * A.k[2][i] + CK::B54 * d[c]);
    } else if (S == 5) {
      A.k[4][i] = d[c];
      A.ynext[i] = y + h * (CK::B61 * A.k[0][i] + CK::B62 * A.k[1][i] + CK::B63 * A.k[2][i] + CK::B64 * A.k[3][i] +
                            CK::B65 * d[c]);
    } else {
      const double k1 = A.k[0][i], k3 = A.k[2][i], k4 = A.k[3][i], k5 = A.k[4][i];
      const double yo = y + h * (CK::C1 * k1 + CK::C3 * k3 + CK::C4 * k4 + CK::C6 * d[c]);
      const double ye = h * (CK::DC1 * k1 + CK::DC3 * k3 + CK::DC4 * k4 + CK::DC5 * k5 + CK::DC6 * d[c]);
      A.ynext[i] = yo;
      double e = fabs(ye / A.yscal[i]);
      if (!(e == e)) e = INFINITY;
      emax = fmax(emax, e);
      if (c == 2 && (yo > 0 || sig_nv > 0)) tensile = 1;  // Newton-V: the last stage's normal stress
    }
  }
}

template <int S>
__global__ void __launch_bounds__(256) stage_kernel(StageArgs A) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  double emax = 0.0;
  int tensile = 0;
  if (k < A.n) stage_body<S>(A, k, A.h, A.t_try, A.tol, A.p.eta_rad, emax, tensile);
  if (S == 6) {
    __shared__ double red[4];
    __shared__ int tred[4];
    double m = wave_max(emax);
    int tf = __any(tensile) ? 1 : 0;
    if ((threadIdx.x & 63) == 0) { red[threadIdx.x >> 6] = m; tred[threadIdx.x >> 6] = tf; }
    __syncthreads();
    if (threadIdx.x == 0) {
      const double bm = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
      // non-negative doubles order like their bit patterns
      atomicMax(reinterpret_cast<unsigned long long *>(&A.status->errmax_bits), __double_as_longlong(bm));
      if (tred[0] | tred[1] | tred[2] | tred[3]) atomicOr(&A.status->tensile, 1);
    }
  }
}

// The same stage over an ensemble: each element takes its network's step size,
// trial step, tolerance and damping; elements of networks outside the attempt
// are left untouched.  At s = 6 the error norm is reduced per network: a wave
// whose elements all belong to one network reduces first, a wave straddling two
// networks updates per element.
template <int S>
__global__ void __launch_bounds__(256) ens_stage_kernel(EnsStageArgs E) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int nb = k < E.a.n ? E.net[k] : -1;
  const bool on = nb >= 0 && E.active[nb];
  double emax = 0.0;
  int tensile = 0;
  if (on) stage_body<S>(E.a, k, E.h[nb], E.t_try[nb], E.tol[nb], E.eta[nb], emax, tensile);
  if (S == 6) {
    const int first = __shfl(nb, 0, 64);
    const bool uniform = __all(nb == first);
    if (uniform) {
      const double m = wave_max(emax);
      const int tf = __any(tensile) ? 1 : 0;
      if ((threadIdx.x & 63) == 0 && first >= 0 && E.active[first]) {
        atomicMax(reinterpret_cast<unsigned long long *>(&E.status[first].errmax_bits), __double_as_longlong(m));
        if (tf) atomicOr(&E.status[first].tensile, 1);
      }
    } else if (on) {
      atomicMax(reinterpret_cast<unsigned long long *>(&E.status[nb].errmax_bits), __double_as_longlong(emax));
      if (tensile) atomicOr(&E.status[nb].tensile, 1);
    }
  }
}

// ytemp2 = y + B21*h*k1 with each active network's h (ode_bs.f90:213)
__global__ void __launch_bounds__(256) ens_restart_stage2_kernel(EnsStageArgs E) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= E.a.n) return;
  const int nb = E.net[k];
  if (!E.active[nb]) return;
  const double h = E.h[nb];
  const int n = E.a.n;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const int i = c * n + k;
    E.a.ynext[i] = E.a.y[i] + CK::B21 * h * E.a.k[0][i];
  }
}

__global__ void __launch_bounds__(256) monitor_kernel(const double *__restrict__ Vold, const double *__restrict__ Vnew,
                                                      double h, const int *__restrict__ fault_of, int n,
                                                      double *__restrict__ slip, unsigned long long *__restrict__ fvmax) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int f = k < n ? fault_of[k] : -1;
  double v = 0.0;
  if (f >= 0) {
    slip[k] = slip[k] + 0.5 * (Vold[k] + Vnew[k]) * h;
    v = fabs(Vnew[k]);
  }
  const int first = __shfl(f, 0, 64);
  if (__all(f == first)) {
    const double m = wave_max(v);
    if ((threadIdx.x & 63) == 0 && first >= 0) atomicMax(&fvmax[first], (unsigned long long)__double_as_longlong(m));
  } else if (f >= 0) {
    atomicMax(&fvmax[f], (unsigned long long)__double_as_longlong(v));
  }
}

__global__ void __launch_bounds__(64) fill_u32_kernel(unsigned *p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

__global__ void __launch_bounds__(256) fill_u64_kernel(unsigned long long *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

__global__ void __launch_bounds__(256) ens_clear_status_kernel(StatusWord *st, const int *active, int nnet) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nnet && (!active || active[b])) { st[b].errmax_bits = 0ull; st[b].tensile = 0; }
}

__global__ void __launch_bounds__(256) ens_cg_dt_kernel(CgSys *sys, int nnet, int nslot, const double *t_try,
                                                        const double *h, int first_slot) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nnet * nslot) return;
  const int b = sys[q].net, slot = first_slot + q % nslot;  // nslot consecutive systems per network
  const double A[6] = {0.0, CK::A2, CK::A3, CK::A4, CK::A5, CK::A6};
  sys[q].dt = slot == 0 ? t_try[b] : 0.0 + A[slot] * h[b];
}

// ytemp2 = y + B21*h*k1 for a retried step size (ode_bs.f90:213)
__global__ void __launch_bounds__(256) restart_stage2_kernel(const double *__restrict__ y, const double *__restrict__ k1,
                                                             double h, int n3, double *__restrict__ yt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3) yt[i] = y[i] + CK::B21 * h * k1[i];
}

// ---------------------------------------------------------------------------
// compute_P (src/fluid_diffusion.f90:24-405): one workgroup per time step dt.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_sum(double v, double *red) {
  v = wave_sum(v);
  const int wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  for (int q = 0; q < nw; q++) s += red[q];
  return s;
}

__device__ __forceinline__ double block_max(double v, double *red) {
  v = wave_max(v);
  const int wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  for (int q = 0; q < nw; q++) s = fmax(s, red[q]);
  return s;
}

// Per network (one 1024-thread workgroup each, in a fixed order): over its real
// elements, max |V|, the mean of log10 |V| and the fraction with |V| > 1e-3 m/s;
// out[3 b + 0..2].
__global__ void __launch_bounds__(1024) ens_vstats_kernel(const double *__restrict__ V,
                                                         const uint8_t *__restrict__ fake,
                                                         const CgNet *__restrict__ nets, double *__restrict__ out) {
  __shared__ double red[16];
  const CgNet c = nets[blockIdx.x];
  double mx = 0.0, sl = 0.0, fast = 0.0, nreal = 0.0;
  for (int k = threadIdx.x; k < c.n; k += blockDim.x) {
    const int e = c.base + k;
    if (fake[e]) continue;
    const double v = fabs(V[e]);
    mx = fmax(mx, v);
    sl += log10(fmax(v, 1e-30));
    fast += v > 1e-3 ? 1.0 : 0.0;
    nreal += 1.0;
  }
  mx = block_max(mx, red);
  sl = block_sum(sl, red);
  fast = block_sum(fast, red);
  nreal = block_sum(nreal, red);
  if (threadIdx.x == 0) {
    out[3 * blockIdx.x] = mx;
    out[3 * blockIdx.x + 1] = sl / fmax(nreal, 1.0);
    out[3 * blockIdx.x + 2] = fast / fmax(nreal, 1.0);
  }
}

__device__ __forceinline__ double hmean(double pa, double pb) { return 2.0 * pa * pb / (pa + pb); }

// Multi-network launches: restricts the arguments to the system's network
// (offset arrays, its sources, crossings and constants); false when the network
// sits out this launch.
__device__ __forceinline__ bool cg_view(CgArgs &A, CrossRows &X, const CgSys &S) {
  if (!A.nets) return true;
  if (A.active && !A.active[S.net]) return false;
  const CgNet c = A.nets[S.net];
  A.n = c.n;
  A.ds += c.base;
  A.P += c.base;
  A.perm += c.base;
  A.perm_diag += c.base;
  A.perm_x += c.xbase;
  A.perm_x_diag += c.xbase;
  A.cfac = c.cfac;
  A.time = c.time;
  A.tol = c.tol;
  A.por = c.por;
  A.beta = c.beta;
  A.src_elem += c.src0;
  A.src_Q += c.src0;
  A.src_tb += c.src0;
  A.src_te += c.src0;
  A.nsrc = c.nsrc;
  A.cross += 6 * (int64_t)c.cross0;
  A.ncross = c.ncross;
  if (X.start) X.start += c.base;
  return true;
}

// M v for system s: rows (A1 v_{k-1} + A2 v_k + A3 v_{k+1}) / diag_k plus the crossing terms.
// out = (D^-1 M) v; with dot != nullptr and no crossing terms also this thread's
// share of v.out, summed in the order of the caller's own loop over k
__device__ void cg_matvec(const CgArgs &A, const CgSys &S, double dt, const double *v, double *out,
                          double *dot = nullptr) {
  const int n = A.n;
  double part = 0.0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    double r;
    if (k == 0) r = S.A2[0] * v[0] + S.A3[0] * v[1];
    else if (k == n - 1) r = S.A1[k] * v[k - 1] + S.A2[k] * v[k];
    else r = S.A1[k] * v[k - 1] + S.A2[k] * v[k] + S.A3[k] * v[k + 1];
    const double o = r / S.diag[k];
    out[k] = o;
    part += v[k] * o;
  }
  if (dot && A.ncross == 0) *dot = part;
  if (A.ncross > 0) {
    __syncthreads();
    if (threadIdx.x == 0) {
      const double *p = A.perm, *ds = A.ds;
      const int kq = n > 2 ? n : 2;
      const double dq = A.cfac * (ds[kq - 2] * ds[kq - 2]);
      for (int q = 0; q < A.ncross; q++) {
        const int64_t n1 = A.cross[6 * q + 2], n2 = A.cross[6 * q + 3];  // 1-based node ids
        // element e (1-based) -> index e-1
        double kl, kr, a1, a2, a3;
        kl = hmean(p[n1 - 2], p[n2 - 2]); kr = hmean(p[n1 - 2], p[n2 - 1]);
        a1 = -(kl * dt / dq); a3 = -(kr * dt / dq); a2 = kl * dt / dq + kr * dt / dq;
        out[n1 - 2] = out[n1 - 2] + (a1 * v[n2 - 2] + a2 * v[n1 - 2] + a3 * v[n2 - 1]) / S.diag[n1 - 2];
        kl = hmean(p[n1 - 1], p[n2 - 2]); kr = hmean(p[n1 - 1], p[n2 - 1]);
        a1 = -(kl * dt / dq); a3 = -(kr * dt / dq); a2 = kl * dt / dq + kr * dt / dq;
        out[n1 - 1] = out[n1 - 1] + (a1 * v[n2 - 2] + a2 * v[n1 - 1] + a3 * v[n2 - 1]) / (double)n1;
        kl = hmean(p[n1 - 2], p[n2 - 2]); kr = hmean(p[n1 - 1], p[n2 - 2]);
        a1 = -(kl * dt / dq); a3 = -(kr * dt / dq); a2 = kl * dt / dq + kr * dt / dq;
        out[n2 - 2] = out[n2 - 2] + (a1 * v[n1 - 2] + a2 * v[n2 - 2] + a3 * v[n1 - 1]) / S.diag[n2 - 2];
        kl = hmean(p[n1 - 2], p[n2 - 1]); kr = hmean(p[n1 - 1], p[n2 - 1]);
        a1 = -(kl * dt / dq); a3 = -(kr * dt / dq); a2 = kl * dt / dq + kr * dt / dq;
        out[n2 - 1] = out[n2 - 1] + (a1 * v[n1 - 2] + a2 * v[n2 - 1] + a3 * v[n1 - 1]) / S.diag[n2 - 1];
      }
    }
  }
  __syncthreads();
}

__global__ void __launch_bounds__(1024) cg_kernel(CgArgs A) {
  __shared__ double red[16];
  const CgSys S = A.sys[blockIdx.x];
  CrossRows nox{nullptr, nullptr};
  if (!cg_view(A, nox, S)) return;
  const double dt = S.dt;
  const int n = A.n;
  const double c = A.cfac;
  const double *kx = A.perm_x;  // 0-based: kx[k] is permeability_x(k+1)
  const double *ds = A.ds;
  // diag (compute_diag :308-405) with the permeability_x in effect before this call
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    double d;
    if (k == 0) d = 1.0 + A.perm_x_diag[1] * dt / (c * (ds[0] * ds[0]));
    else if (k == n - 1) d = 1.0 + A.perm_x_diag[n - 1] * dt / (c * (ds[n - 1] * ds[n - 1]));
    else d = 1.0 + (A.perm_x_diag[k] + A.perm_x_diag[k + 1]) * dt / (c * (ds[k] * ds[k]));
    S.diag[k] = d;
    // stencil coefficients of compute_MatrixVector (:175-204) with the current permeability_x
    if (k == 0) {
      S.A2[0] = 1.0 + kx[1] * dt / (c * (ds[0] * ds[0]));
      S.A3[0] = -(kx[1] * dt / (c * (ds[0] * ds[0])));
    } else if (k == n - 1) {
      S.A1[k] = -(kx[n - 1] * dt / (c * (ds[n - 1] * ds[n - 1])));
      S.A2[k] = 1.0 + kx[n - 1] * dt / (c * (ds[n - 1] * ds[n - 1]));
    } else {
      S.A1[k] = -(kx[k] * dt / (c * (ds[k - 1] * ds[k - 1])));
      S.A2[k] = 1.0 + (kx[k] + kx[k + 1]) * dt / (c * (ds[k] * ds[k]));
      S.A3[k] = -(kx[k + 1] * dt / (c * (ds[k + 1] * ds[k + 1])));
    }
  }
  __syncthreads();
  if (A.ncross > 0 && threadIdx.x == 0) {
    const double *p = A.perm_diag;
    const double d1 = c * (ds[0] * ds[0]);
    for (int q = 0; q < A.ncross; q++) {
      const int64_t n1 = A.cross[6 * q + 2], n2 = A.cross[6 * q + 3];
      const int64_t e[4] = {n1 - 2, n1 - 1, n2 - 2, n2 - 1};
      const double kls[4] = {hmean(p[n1 - 2], p[n2 - 2]), hmean(p[n1 - 1], p[n2 - 2]), hmean(p[n1 - 2], p[n2 - 2]),
                             hmean(p[n1 - 2], p[n2 - 1])};
      const double krs[4] = {hmean(p[n1 - 2], p[n2 - 1]), hmean(p[n1 - 1], p[n2 - 1]), hmean(p[n1 - 1], p[n2 - 2]),
                             hmean(p[n1 - 1], p[n2 - 1])};
      for (int z = 0; z < 4; z++) S.diag[e[z]] = S.diag[e[z]] + kls[z] * dt / d1 + krs[z] * dt / d1;
    }
  }
  __syncthreads();
  // right-hand side b1 = P / diag (+ sources, :52-64) and the initial residual
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    S.b[k] = A.P[k] / S.diag[k];
    S.x[k] = A.P[k];
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 0; q < A.nsrc; q++) {
      const double tt = A.time + dt;
      if (tt >= A.src_tb[q] && tt < A.src_te[q]) {
        const int64_t e = A.src_elem[q] - 1;
        S.b[e] = S.b[e] + (dt * A.src_Q[q] / (A.por * A.beta) / ds[e]) / S.diag[e];
      }
    }
  __syncthreads();
  cg_matvec(A, S, dt, S.x, S.Ap);
  double emax = 0.0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const double r = S.b[k] - S.Ap[k];
    S.r[k] = r;
    S.p[k] = r;
    emax = fmax(emax, fabs(r / 1.0));
  }
  double errmax = block_max(emax, red) / A.tol;
  double rr = 0.0;
  {
    double part = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) part += S.r[k] * S.r[k];
    rr = block_sum(part, red);
  }
  int it = 0, ier = 0;
  while (errmax >= 1.0) {
    it++;
    // p.Ap in the same pass as A p when no crossing term changes A p afterwards
    double part = 0.0;
    cg_matvec(A, S, dt, S.p, S.Ap, &part);
    if (A.ncross > 0) {
      part = 0.0;
      for (int k = threadIdx.x; k < n; k += blockDim.x) part += S.p[k] * S.Ap[k];
    }
    const double pAp = block_sum(part, red);
    const double alpha = rr / pAp;
    double p1 = 0.0, em = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
      S.x[k] = S.x[k] + alpha * S.p[k];
      const double r1 = S.r[k] - alpha * S.Ap[k];
      S.r[k] = r1;
      p1 += r1 * r1;
      em = fmax(em, fabs(r1 / 1.0));
    }
    const double r1r1 = block_sum(p1, red);
    const double beta = r1r1 / rr;
    for (int k = threadIdx.x; k < n; k += blockDim.x) S.p[k] = S.r[k] + beta * S.p[k];
    rr = r1r1;
    if (it > kCgMaxIter) { ier = 1; break; }
    errmax = block_max(em, red) / A.tol;
  }
  double xmax = 0.0;
  if (!ier) {
    const bool small = dt < (double)1e-9f;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
      const double xk = S.x[k];
      S.Pt[k] = xk;
      S.Pd[k] = small ? 0.0 : (xk - A.P[k]) / dt;
      xmax = fmax(xmax, fabs(xk));
    }
  }
  xmax = block_max(xmax, red);
  if (threadIdx.x == 0) {
    if (!ier && xmax > 1.7976931348623157e308) ier = 1;
    S.result[0] = ier;
    S.result[1] = it;
  }
}

// Cooperative compute_P: every system (time step dt) is spread over `wps`
// workgroups of 256 threads, kCgEPT elements per thread held in registers (x, r,
// p, A p and the stencil coefficients).  Neighbour and crossing values of p go
// through global memory; the dot products are per-workgroup partials summed in
// a fixed order after each grid barrier, so every workgroup of a system computes
// the same alpha, beta and stopping decision.  Three grid barriers per iteration.
__device__ __forceinline__ void cross_ends(int z, int64_t n1, int64_t n2, int64_t &ea, int64_t &eb, int64_t &ec,
                                           int64_t &pl0, int64_t &pl1, int64_t &pr0, int64_t &pr1) {
  // for slot z: the row's A1/A2/A3 operands (0-based elements) and the permeability pairs
  // of the left and right harmonic means (src/fluid_diffusion.f90:221-291, :349-401)
  switch (z) {
    case 0: ea = n2 - 2; eb = n1 - 2; ec = n2 - 1; pl0 = n1 - 2; pl1 = n2 - 2; pr0 = n1 - 2; pr1 = n2 - 1; break;
    case 1: ea = n2 - 2; eb = n1 - 1; ec = n2 - 1; pl0 = n1 - 1; pl1 = n2 - 2; pr0 = n1 - 1; pr1 = n2 - 1; break;
    case 2: ea = n1 - 2; eb = n2 - 2; ec = n1 - 1; pl0 = n1 - 2; pl1 = n2 - 2; pr0 = n1 - 1; pr1 = n2 - 2; break;
    default: ea = n1 - 2; eb = n2 - 1; ec = n1 - 1; pl0 = n1 - 2; pl1 = n2 - 1; pr0 = n1 - 1; pr1 = n2 - 1; break;
  }
}

// Values of a vector at other elements: either read directly (v1 == nullptr) or
// formed as v0[e] + beta * v1[e] (the new search direction from the previous
// residual and direction, as their owner forms it).
struct NbVec {
  const double *v0, *v1;
  double beta;
  __device__ __forceinline__ double operator()(int64_t e) const { return v1 ? v0[e] + beta * v1[e] : v0[e]; }
};

// v at other elements, read from a workgroup-shared copy
struct LdsVec {
  const double *v;
  __device__ __forceinline__ double operator()(int64_t e) const { return v[e]; }
};

// M v at element k (row-scaled stencil plus crossing terms)
template <class Vec>
__device__ __forceinline__ double coop_mv(const CgArgs &A, const CrossRows &X, int k, int t0, int t1, double vk,
                                          const Vec &v, double a1, double a2, double a3, double dg, double dt,
                                          double dq) {
  const int n = A.n;
  double r;
  if (k == 0) r = a2 * vk + a3 * v(1);
  else if (k == n - 1) r = a1 * v(k - 1) + a2 * vk;
  else r = a1 * v(k - 1) + a2 * vk + a3 * v(k + 1);
  r = r / dg;
  for (int t = t0; t < t1; t++) {
    const int cidx = X.ent[2 * t], z = X.ent[2 * t + 1];
    const int64_t n1 = A.cross[6 * cidx + 2], n2 = A.cross[6 * cidx + 3];
    int64_t ea, eb, ec, pl0, pl1, pr0, pr1;
    cross_ends(z, n1, n2, ea, eb, ec, pl0, pl1, pr0, pr1);
    const double kl = hmean(A.perm[pl0], A.perm[pl1]), kr = hmean(A.perm[pr0], A.perm[pr1]);
    const double c1 = -(kl * dt / dq), c3 = -(kr * dt / dq), c2 = kl * dt / dq + kr * dt / dq;
    const double vb = (eb == k) ? vk : v(eb);
    const double num = c1 * v(ea) + c2 * vb + c3 * v(ec);
    r = r + num / (z == 1 ? (double)n1 : dg);
  }
  return r;
}


// Messages between the workgroups of one system of a cooperative compute_P.  A
// double travels as two 64-bit words, each holding 32 bits of it beside the
// 32-bit epoch of the message; a reader takes the value once both words carry
// the epoch it waits for.  Every word is written and read with relaxed
// device-scope atomics, so each is seen whole and eventually, and no fence or
// grid barrier is needed: each word's epoch vouches for its own half.  A wait
// that outlasts kCgSpins polls gives up and sets `late`, so a fault can never
// hang the device.
constexpr int kCgSpins = 1 << 18;

__device__ __forceinline__ void ll_send(unsigned long long *w0, unsigned long long *w1, double v, unsigned epoch) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned long long e = (unsigned long long)epoch << 32;
  __hip_atomic_store(w0, e | (b & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(w1, e | (b >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double ll_recv(unsigned long long *w0, unsigned long long *w1, unsigned epoch,
                                          bool &late) {
  for (int spins = 0;; spins++) {
    const unsigned long long a = __hip_atomic_load(w0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long b = __hip_atomic_load(w1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((unsigned)(a >> 32) == epoch && (unsigned)(b >> 32) == epoch)
      return __longlong_as_double((long long)((b << 32) | (a & 0xffffffffull)));
    if (spins >= kCgSpins) { late = true; return 0.0; }
    __builtin_amdgcn_s_sleep(1);
  }
}

// Workgroup records of a system: 8 words per workgroup, [0,1] the p.Ap partial
// (or the final overflow flag), [2,3] r.r, [4,5] max|r|.
constexpr int kCgRec = kCgRecordWords;

// Sum (or max) over the wps workgroups' messages at word `off` of their
// records, one per thread, then a fixed tree: every workgroup of the system
// gets the same double.
__device__ __forceinline__ double ll_sum(unsigned long long *rec, int wps, int off, unsigned epoch, double *red,
                                         bool &late) {
  double v = 0.0;
  for (int q = threadIdx.x; q < wps; q += blockDim.x)
    v += ll_recv(rec + kCgRec * q + off, rec + kCgRec * q + off + 1, epoch, late);
  return block_sum(v, red);
}

// Sum and max over the workgroup in one pass (red holds 16 doubles).
__device__ __forceinline__ void block_sum_max(double &s, double &m, double *red) {
  wave_sum_max_full(s, m);
  const int wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) { red[wid] = s; red[8 + wid] = m; }
  __syncthreads();
  double a = 0.0, b = 0.0;
  for (int q = 0; q < nw; q++) { a += red[q]; b = fmax(b, red[8 + q]); }
  s = a;
  m = b;
}

// The sum of the messages at word off_s and the max of those at off_m.
__device__ __forceinline__ void ll_sum_max(unsigned long long *rec, int wps, int off_s, int off_m, unsigned epoch,
                                           double *red, bool &late, double &s, double &m) {
  s = 0.0;
  m = 0.0;
  for (int q = threadIdx.x; q < wps; q += blockDim.x) {
    s += ll_recv(rec + kCgRec * q + off_s, rec + kCgRec * q + off_s + 1, epoch, late);
    m = fmax(m, ll_recv(rec + kCgRec * q + off_m, rec + kCgRec * q + off_m + 1, epoch, late));
  }
  block_sum_max(s, m, red);
}

// The crossing entries of one workgroup's span of rows, staged in LDS at the
// start of a cooperative compute_P.  Entry j (in row order) contributes
// (c1 v(ea) + c2 v(eb) + c3 v(ec)) / dv to its row (src/fluid_diffusion.f90:221-291);
// term[j] holds that contribution for the current direction, pa/pb/pc[j] the
// direction at ea/eb/ec when that element lies outside the span.
struct SpanEntries {
  double *c1, *c2, *c3, *dv, *term, *pa, *pb, *pc;
  int *ea, *eb, *ec;
  __device__ SpanEntries(double *base, int cap)
      : c1(base), c2(base + cap), c3(base + 2 * cap), dv(base + 3 * cap), term(base + 4 * cap),
        pa(base + 5 * cap), pb(base + 6 * cap), pc(base + 7 * cap),
        ea(reinterpret_cast<int *>(base + 8 * cap)), eb(ea + cap), ec(ea + 2 * cap) {}
};

// Cooperative compute_P: every system (time step dt) is spread over `wps`
// workgroups of 256 threads, each owning a span of kCgSpan rows, kCgEPT per
// thread held in registers (x, r, p, A p and the stencil coefficients).  Per
// iteration: the span's direction goes to LDS, with the halo values of the
// neighbouring spans and the crossing operands outside the span formed from
// messages; the span's crossing terms are evaluated in parallel; A p row by row;
// the p.Ap partials are exchanged as messages and summed in a fixed order, so
// every workgroup of a system computes the same alpha; x and r are updated, the
// new residual is sent for the elements other workgroups read (the span ends and
// the crossing elements), and the r.r and max|r| partials are exchanged like
// p.Ap (same beta and stopping decision everywhere).  A reader forms the new
// direction at a remote element itself, fma(beta, p_old, r), from the residual
// it received and the direction it kept from the last iteration: the owner
// computes the same fma on the same doubles, so both hold the same direction,
// and an iteration waits on two message exchanges instead of three.  Systems
// never wait for each other: one that converges early leaves the device to the
// others.  Message epochs are seq * 256 + a step number, so no word left by an
// earlier launch can match; a residual for the next step overwrites a word only
// after the p.Ap exchange, which every workgroup of the system joins after its
// last read of that word.  Dynamic LDS: (kCgSpan + 2) doubles for p, then 8
// doubles and 3 ints per crossing entry of the span (cap entries; the launcher
// checks that every span fits).
__global__ void __launch_bounds__(256) cg_coop_kernel(CgArgs A, CrossRows X, int wps, int cap, unsigned seq,
                                                      unsigned long long *records) {
  extern __shared__ double coop_lds[];
  __shared__ double red[16];
  int sys, part, first;  // this workgroup's system, its span index, the system's first workgroup
  if (A.wg_map) {        // systems of several networks, of their own sizes
    const int4 m = A.wg_map[blockIdx.x];
    sys = m.x; part = m.y; wps = m.z; first = m.w;
  } else {
    sys = blockIdx.x / wps; part = blockIdx.x - sys * wps; first = sys * wps;
  }
  const CgSys S = A.sys[sys];
  if (!cg_view(A, X, S)) return;  // every workgroup of a network that sits out returns
  const int n = A.n;
  const double dt = S.dt, c = A.cfac;
  const double *ds = A.ds, *kx = A.perm_x, *kxd = A.perm_x_diag, *P = A.P;
  const int span0 = part * kCgSpan, span1 = min(span0 + kCgSpan, n);
  const int base = span0 + threadIdx.x;
  const int kq = n > 2 ? n : 2;
  const double dq = c * (ds[kq - 2] * ds[kq - 2]);
  const double d1 = c * (ds[0] * ds[0]);
  // direction messages, one per element: the two words of element e are w0[e], w1[e]
  unsigned long long *w0 = reinterpret_cast<unsigned long long *>(S.p);
  unsigned long long *w1 = reinterpret_cast<unsigned long long *>(S.Ap);
  unsigned long long *rec = records + (size_t)kCgRec * first;
  unsigned long long *myrec = rec + (size_t)kCgRec * part;
  const unsigned ep0 = seq << 8;
  double *lp = coop_lds;  // lp[1 + k - span0] = p at element k; lp[0] and lp[1 + span1 - span0] are the halo
  const SpanEntries E(coop_lds + kCgSpan + 2, cap);
  const int t_lo = X.start[span0], nent = X.start[span1] - t_lo;
  double a1[kCgEPT], a2[kCgEPT], a3[kCgEPT], dg[kCgEPT], x[kCgEPT], r[kCgEPT], p[kCgEPT], ap[kCgEPT];
  int xs[kCgEPT], xe[kCgEPT];  // this element's range of crossing entries
  unsigned shared = 0;         // bit q: element q is read by other workgroups
  double part_rr = 0.0, part_em = 0.0;
#pragma unroll
  for (int q = 0; q < kCgEPT; q++) {
    const int k = base + q * 256;
    a1[q] = 0.0; a2[q] = 0.0; a3[q] = 0.0; dg[q] = 1.0; x[q] = 0.0; r[q] = 0.0; p[q] = 0.0; ap[q] = 0.0;
    xs[q] = 0; xe[q] = 0;
    if (k < n) {
      xs[q] = X.start[k];
      xe[q] = X.start[k + 1];
      if (k == span0 || k == span1 - 1 || xs[q] < xe[q]) shared |= 1u << q;
      // compute_diag (:308-405) with the permeability in effect before this call
      double d;
      if (k == 0) d = 1.0 + kxd[1] * dt / (c * (ds[0] * ds[0]));
      else if (k == n - 1) d = 1.0 + kxd[n - 1] * dt / (c * (ds[n - 1] * ds[n - 1]));
      else d = 1.0 + (kxd[k] + kxd[k + 1]) * dt / (c * (ds[k] * ds[k]));
      for (int t = xs[q]; t < xe[q]; t++) {
        const int cidx = X.ent[2 * t], z = X.ent[2 * t + 1];
        const int64_t n1 = A.cross[6 * cidx + 2], n2 = A.cross[6 * cidx + 3];
        int64_t ea, eb, ec, pl0, pl1, pr0, pr1;
        cross_ends(z, n1, n2, ea, eb, ec, pl0, pl1, pr0, pr1);
        const double kl = hmean(A.perm_diag[pl0], A.perm_diag[pl1]), kr = hmean(A.perm_diag[pr0], A.perm_diag[pr1]);
        d = d + kl * dt / d1 + kr * dt / d1;
      }
      dg[q] = d;
      // the crossing entries of this row with this call's coefficients (:221-291)
      for (int t = xs[q]; t < xe[q]; t++) {
        const int cidx = X.ent[2 * t], z = X.ent[2 * t + 1];
        const int64_t n1 = A.cross[6 * cidx + 2], n2 = A.cross[6 * cidx + 3];
        int64_t ea, eb, ec, pl0, pl1, pr0, pr1;
        cross_ends(z, n1, n2, ea, eb, ec, pl0, pl1, pr0, pr1);
        const double kl = hmean(A.perm[pl0], A.perm[pl1]), kr = hmean(A.perm[pr0], A.perm[pr1]);
        const int j = t - t_lo;
        E.c1[j] = -(kl * dt / dq);
        E.c3[j] = -(kr * dt / dq);
        E.c2[j] = kl * dt / dq + kr * dt / dq;
        E.dv[j] = z == 1 ? (double)n1 : d;
        E.ea[j] = (int)ea;
        E.eb[j] = (int)eb;
        E.ec[j] = (int)ec;
      }
      // compute_MatrixVector coefficients (:175-204)
      if (k == 0) {
        a2[q] = 1.0 + kx[1] * dt / (c * (ds[0] * ds[0]));
        a3[q] = -(kx[1] * dt / (c * (ds[0] * ds[0])));
      } else if (k == n - 1) {
        a1[q] = -(kx[n - 1] * dt / (c * (ds[n - 1] * ds[n - 1])));
        a2[q] = 1.0 + kx[n - 1] * dt / (c * (ds[n - 1] * ds[n - 1]));
      } else {
        a1[q] = -(kx[k] * dt / (c * (ds[k - 1] * ds[k - 1])));
        a2[q] = 1.0 + (kx[k] + kx[k + 1]) * dt / (c * (ds[k] * ds[k]));
        a3[q] = -(kx[k + 1] * dt / (c * (ds[k + 1] * ds[k + 1])));
      }
      // b1 = P/diag plus the active sources (:45-64)
      double b = P[k] / d;
      for (int s2 = 0; s2 < A.nsrc; s2++) {
        const double tt = A.time + dt;
        if (A.src_elem[s2] - 1 == k && tt >= A.src_tb[s2] && tt < A.src_te[s2])
          b = b + (dt * A.src_Q[s2] / (A.por * A.beta) / ds[k]) / d;
      }
      x[q] = P[k];
      const NbVec vP{P, nullptr, 0.0};
      const double mv = coop_mv(A, X, k, xs[q], xe[q], P[k], vP, a1[q], a2[q], a3[q], d, dt, dq);
      r[q] = b - mv;
      p[q] = r[q];
      if (shared & (1u << q)) ll_send(w0 + k, w1 + k, r[q], ep0 + 1);  // r_0, which is p_1
      part_rr += r[q] * r[q];
      part_em = fmax(part_em, fabs(r[q] / 1.0));
    }
  }
  bool late = false;
  block_sum_max(part_rr, part_em, red);
  if (threadIdx.x == 0) {
    ll_send(myrec + 2, myrec + 3, part_rr, ep0);
    ll_send(myrec + 4, myrec + 5, part_em, ep0);
  }
  double rr, em;
  ll_sum_max(rec, wps, 2, 4, ep0, red, late, rr, em);
  bool timed_out = __syncthreads_or(late);
  bool fail = false;
  int it = 0;
  double beta = 0.0;      // of the last iteration
  double halo_old = 0.0;  // thread 0: p at span0 - 1; the last thread: p at span1
  // the direction at a remote element from its residual message (p_1 = r_0)
  const auto remote_p = [&](double rv, double p_old) { return it == 0 ? rv : fma(beta, p_old, rv); };
  // every workgroup of the system receives the same partials and sums them in the
  // same order, so they all take the same decisions
  while (em / A.tol >= 1.0 && !timed_out) {
    const unsigned ep = ep0 + (unsigned)it + 1;  // r_it for p_{it+1}, and this iteration's partials
#pragma unroll
    for (int q = 0; q < kCgEPT; q++) {
      const int k = base + q * 256;
      if (k < n) lp[1 + k - span0] = p[q];
    }
    if (threadIdx.x == 0 && span0 > 0) {
      halo_old = remote_p(ll_recv(w0 + span0 - 1, w1 + span0 - 1, ep, late), halo_old);
      lp[0] = halo_old;
    }
    if (threadIdx.x == blockDim.x - 1 && span1 < n) {
      halo_old = remote_p(ll_recv(w0 + span1, w1 + span1, ep, late), halo_old);
      lp[1 + span1 - span0] = halo_old;
    }
    __syncthreads();  // p of the span and its halo are in LDS
    for (int j = threadIdx.x; j < nent; j += blockDim.x) {
      const int ea = E.ea[j], eb = E.eb[j], ec = E.ec[j];
      double va, vb, vc;
      if (ea >= span0 && ea < span1) va = lp[1 + ea - span0];
      else va = E.pa[j] = remote_p(ll_recv(w0 + ea, w1 + ea, ep, late), E.pa[j]);
      if (eb >= span0 && eb < span1) vb = lp[1 + eb - span0];
      else vb = E.pb[j] = remote_p(ll_recv(w0 + eb, w1 + eb, ep, late), E.pb[j]);
      if (ec >= span0 && ec < span1) vc = lp[1 + ec - span0];
      else vc = E.pc[j] = remote_p(ll_recv(w0 + ec, w1 + ec, ep, late), E.pc[j]);
      const double num = E.c1[j] * va + E.c2[j] * vb + E.c3[j] * vc;
      E.term[j] = num / E.dv[j];
    }
    __syncthreads();  // the crossing terms are in LDS
    double part_pap = 0.0;
#pragma unroll
    for (int q = 0; q < kCgEPT; q++) {
      const int k = base + q * 256;
      if (k < n) {
        const int l = 1 + k - span0;
        double v;
        if (k == 0) v = a2[q] * p[q] + a3[q] * lp[l + 1];
        else if (k == n - 1) v = a1[q] * lp[l - 1] + a2[q] * p[q];
        else v = a1[q] * lp[l - 1] + a2[q] * p[q] + a3[q] * lp[l + 1];
        v = v / dg[q];
        for (int t = xs[q]; t < xe[q]; t++) v = v + E.term[t - t_lo];
        ap[q] = v;
        part_pap += p[q] * ap[q];
      }
    }
    part_pap = block_sum(part_pap, red);
    if (threadIdx.x == 0) ll_send(myrec + 0, myrec + 1, part_pap, ep);
    const double pAp = ll_sum(rec, wps, 0, ep, red, late);
    const double alpha = rr / pAp;
    double part_r1 = 0.0, part_m = 0.0;
#pragma unroll
    for (int q = 0; q < kCgEPT; q++) {
      const int k = base + q * 256;
      if (k < n) {
        x[q] = x[q] + alpha * p[q];
        r[q] = r[q] - alpha * ap[q];
        part_r1 += r[q] * r[q];
        part_m = fmax(part_m, fabs(r[q] / 1.0));
        if (shared & (1u << q)) ll_send(w0 + k, w1 + k, r[q], ep + 1);
      }
    }
    block_sum_max(part_r1, part_m, red);
    if (threadIdx.x == 0) {
      ll_send(myrec + 2, myrec + 3, part_r1, ep);
      ll_send(myrec + 4, myrec + 5, part_m, ep);
    }
    double r1r1;
    ll_sum_max(rec, wps, 2, 4, ep, red, late, r1r1, em);
    beta = r1r1 / rr;
#pragma unroll
    for (int q = 0; q < kCgEPT; q++) p[q] = fma(beta, p[q], r[q]);  // as remote_p
    rr = r1r1;
    it++;
    timed_out = __syncthreads_or(late);
    if (it > kCgMaxIter) { fail = true; break; }
  }
  // test hook: the host makes one workgroup of every system report a lost message
  if (A.late_part == part + 1) timed_out = true;
  int bad = 0;
  if (!fail) {
    const bool small = dt < (double)1e-9f;
#pragma unroll
    for (int q = 0; q < kCgEPT; q++) {
      const int k = base + q * 256;
      if (k < n) {
        S.Pt[k] = x[q];
        S.Pd[k] = small ? 0.0 : (x[q] - P[k]) / dt;
        if (fabs(x[q]) > 1.7976931348623157e308) bad = 1;
      }
    }
  }
  // every workgroup's flag goes to the system's first workgroup, which reports:
  // 2 if any workgroup waited in vain for a message (its results are not to be
  // trusted), else 1 if any x overflowed or the solve failed
  const double b2 = block_max(bad ? 1.0 : 0.0, red);
  const double flag = timed_out ? 2.0 : b2;
  if (threadIdx.x == 0) ll_send(myrec + 0, myrec + 1, flag, ep0 + 255);
  if (part == 0) {
    double nflag, maxflag;
    ll_sum_max(rec, wps, 0, 0, ep0 + 255, red, late, nflag, maxflag);
    timed_out = __syncthreads_or(late) || maxflag >= 2.0;
    if (threadIdx.x == 0) {
      S.result[0] = timed_out ? 2 : ((fail || maxflag > 0.0) ? 1 : 0);
      S.result[1] = it;
    }
  }
}

// compute_P for systems that fit one workgroup (n <= kCgLdsThreads * EPT): one
// workgroup per system, EPT elements per thread in registers (interleaved,
// k = q * kCgLdsThreads + tid), the search direction mirrored in LDS for the
// neighbour and crossing reads.  Same recurrences, stopping rule and failure
// rule as cg_coop_kernel; three workgroup barriers per iteration and no grid
// barrier, which is what makes it the faster path for faults of a few thousand
// elements.  The row scaling is folded into the stencil coefficients
// (A1/diag, A2/diag, A3/diag), so a row costs three multiply-adds and no divide.
template <class Vec>
__device__ __forceinline__ double lds_mv(const CgArgs &A, const CrossRows &X, int k, int t0, int t1, double vk,
                                         const Vec &v, double s1, double s2, double s3, double dg, double dt,
                                         double dq) {
  const int km = k > 0 ? k - 1 : k, kp = k < A.n - 1 ? k + 1 : k;  // end rows have a zero coefficient there
  double r = s1 * v(km) + s2 * vk + s3 * v(kp);
  for (int t = t0; t < t1; t++) {
    const int cidx = X.ent[2 * t], z = X.ent[2 * t + 1];
    const int64_t n1 = A.cross[6 * cidx + 2], n2 = A.cross[6 * cidx + 3];
    int64_t ea, eb, ec, pl0, pl1, pr0, pr1;
    cross_ends(z, n1, n2, ea, eb, ec, pl0, pl1, pr0, pr1);
    const double kl = hmean(A.perm[pl0], A.perm[pl1]), kr = hmean(A.perm[pr0], A.perm[pr1]);
    const double c1 = -(kl * dt / dq), c3 = -(kr * dt / dq), c2 = kl * dt / dq + kr * dt / dq;
    const double vb = (eb == k) ? vk : v(eb);
    const double num = c1 * v(ea) + c2 * vb + c3 * v(ec);
    r = r + num / (z == 1 ? (double)n1 : dg);
  }
  return r;
}

template <int EPT>
__global__ void __launch_bounds__(kCgLdsThreads) cg_lds_kernel(CgArgs A, CrossRows X) {
  extern __shared__ double lds_p[];  // [n]
  __shared__ double red1[kCgLdsThreads / 64], red2[kCgLdsThreads / 64], red3[kCgLdsThreads / 64];
  const CgSys S = A.sys[blockIdx.x];
  if (!cg_view(A, X, S)) return;
  const int n = A.n, tid = threadIdx.x, wid = tid >> 6, nw = kCgLdsThreads / 64;
  const double dt = S.dt, c = A.cfac;
  const double *ds = A.ds, *kx = A.perm_x, *kxd = A.perm_x_diag, *P = A.P;
  const int kq = n > 2 ? n : 2;
  const double dq = c * (ds[kq - 2] * ds[kq - 2]);
  const double d1 = c * (ds[0] * ds[0]);
  double a1[EPT], a2[EPT], a3[EPT], dg[EPT], x[EPT], r[EPT], p[EPT];
  int xs[EPT], xe[EPT];  // this element's range of crossing entries
  double part_rr = 0.0, part_em = 0.0;
#pragma unroll
  for (int q = 0; q < EPT; q++) {
    const int k = q * kCgLdsThreads + tid;
    a1[q] = 0.0; a2[q] = 0.0; a3[q] = 0.0; dg[q] = 1.0; x[q] = 0.0; r[q] = 0.0; p[q] = 0.0;
    xs[q] = 0; xe[q] = 0;
    if (k < n) {
      xs[q] = X.start[k];
      xe[q] = X.start[k + 1];
      double d;
      if (k == 0) d = 1.0 + kxd[1] * dt / (c * (ds[0] * ds[0]));
      else if (k == n - 1) d = 1.0 + kxd[n - 1] * dt / (c * (ds[n - 1] * ds[n - 1]));
      else d = 1.0 + (kxd[k] + kxd[k + 1]) * dt / (c * (ds[k] * ds[k]));
      for (int t = xs[q]; t < xe[q]; t++) {
        const int cidx = X.ent[2 * t], z = X.ent[2 * t + 1];
        const int64_t n1 = A.cross[6 * cidx + 2], n2 = A.cross[6 * cidx + 3];
        int64_t ea, eb, ec, pl0, pl1, pr0, pr1;
        cross_ends(z, n1, n2, ea, eb, ec, pl0, pl1, pr0, pr1);
        const double kl = hmean(A.perm_diag[pl0], A.perm_diag[pl1]), kr = hmean(A.perm_diag[pr0], A.perm_diag[pr1]);
        d = d + kl * dt / d1 + kr * dt / d1;
      }
      dg[q] = d;
      if (k == 0) {
        a2[q] = 1.0 + kx[1] * dt / (c * (ds[0] * ds[0]));
        a3[q] = -(kx[1] * dt / (c * (ds[0] * ds[0])));
      } else if (k == n - 1) {
        a1[q] = -(kx[n - 1] * dt / (c * (ds[n - 1] * ds[n - 1])));
        a2[q] = 1.0 + kx[n - 1] * dt / (c * (ds[n - 1] * ds[n - 1]));
      } else {
        a1[q] = -(kx[k] * dt / (c * (ds[k - 1] * ds[k - 1])));
        a2[q] = 1.0 + (kx[k] + kx[k + 1]) * dt / (c * (ds[k] * ds[k]));
        a3[q] = -(kx[k + 1] * dt / (c * (ds[k + 1] * ds[k + 1])));
      }
      a1[q] = a1[q] / d;
      a2[q] = a2[q] / d;
      a3[q] = a3[q] / d;
      double b = P[k] / d;
      for (int s2 = 0; s2 < A.nsrc; s2++) {
        const double tt = A.time + dt;
        if (A.src_elem[s2] - 1 == k && tt >= A.src_tb[s2] && tt < A.src_te[s2])
          b = b + (dt * A.src_Q[s2] / (A.por * A.beta) / ds[k]) / d;
      }
      x[q] = P[k];
      const NbVec vP{P, nullptr, 0.0};
      const double mv = lds_mv(A, X, k, xs[q], xe[q], P[k], vP, a1[q], a2[q], a3[q], d, dt, dq);
      r[q] = b - mv;
      p[q] = r[q];
      lds_p[k] = p[q];
      part_rr += r[q] * r[q];
      part_em = fmax(part_em, fabs(r[q] / 1.0));
    }
  }
  wave_sum_max_full(part_rr, part_em);
  if ((tid & 63) == 0) { red2[wid] = part_rr; red3[wid] = part_em; }
  __syncthreads();  // lds_p and the initial partials are complete
  double rr = 0.0, em = 0.0;
  for (int w = 0; w < nw; w++) { rr += red2[w]; em = fmax(em, red3[w]); }
  int it = 0;
  bool fail = false;
  const LdsVec vp{lds_p};
  while (em / A.tol >= 1.0) {
    it++;
    double ap[EPT];
    double part_pap = 0.0;
#pragma unroll
    for (int q = 0; q < EPT; q++) {
      const int k = q * kCgLdsThreads + tid;
      ap[q] = 0.0;
      if (k < n) {
        ap[q] = lds_mv(A, X, k, xs[q], xe[q], p[q], vp, a1[q], a2[q], a3[q], dg[q], dt, dq);
        part_pap += p[q] * ap[q];
      }
    }
    part_pap = wave_sum_full(part_pap);
    if ((tid & 63) == 0) red1[wid] = part_pap;
    __syncthreads();  // all reads of lds_p done; pAp partials complete
    double pAp = 0.0;
    for (int w = 0; w < nw; w++) pAp += red1[w];
    const double alpha = rr / pAp;
    double part_r1 = 0.0, part_m = 0.0;
#pragma unroll
    for (int q = 0; q < EPT; q++) {
      const int k = q * kCgLdsThreads + tid;
      if (k < n) {
        x[q] = x[q] + alpha * p[q];
        r[q] = r[q] - alpha * ap[q];
        part_r1 += r[q] * r[q];
        part_m = fmax(part_m, fabs(r[q] / 1.0));
      }
    }
    wave_sum_max_full(part_r1, part_m);
    if ((tid & 63) == 0) { red2[wid] = part_r1; red3[wid] = part_m; }
    __syncthreads();  // r.r and max|r| partials complete
    double r1r1 = 0.0;
    em = 0.0;
    for (int w = 0; w < nw; w++) { r1r1 += red2[w]; em = fmax(em, red3[w]); }
    const double beta = r1r1 / rr;
#pragma unroll
    for (int q = 0; q < EPT; q++) {
      const int k = q * kCgLdsThreads + tid;
      p[q] = r[q] + beta * p[q];
      if (k < n) lds_p[k] = p[q];
    }
    rr = r1r1;
    if (it > kCgMaxIter) { fail = true; break; }
    __syncthreads();  // the new direction is in lds_p; red2/red3 are read by all
  }
  int bad = 0;
  if (!fail) {
    const bool small = dt < (double)1e-9f;
#pragma unroll
    for (int q = 0; q < EPT; q++) {
      const int k = q * kCgLdsThreads + tid;
      if (k < n) {
        S.Pt[k] = x[q];
        S.Pd[k] = small ? 0.0 : (x[q] - P[k]) / dt;
        if (fabs(x[q]) > 1.7976931348623157e308) bad = 1;
      }
    }
  }
  bad = __syncthreads_or(bad);
  if (tid == 0) {
    S.result[0] = (fail || bad) ? 1 : 0;
    S.result[1] = it;
  }
}

// permeability update of compute_P (:67-75); writes perm and perm_x
__global__ void __launch_bounds__(256) perm_update_kernel(PermArgs A) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= A.n) return;
  if (A.coupling == 1)
    A.perm[k] = (A.perm_star[k] - A.kmin[k]) * exp(-fabs(A.sigmaN[k] + A.P[k]) / fabs(A.Snk[k])) + A.kmin[k];
}

__global__ void __launch_bounds__(256) perm_x_kernel(const double *__restrict__ perm, double *__restrict__ perm_x, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= 1 && k < n) perm_x[k] = 2.0 * perm[k - 1] * perm[k] / (perm[k - 1] + perm[k]);
}

#pragma clang fp contract(on)

// ---------------------------------------------------------------------------
// Launch helpers (called from the host engine)
// ---------------------------------------------------------------------------
static inline int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

void launch_gather(const double *x, const int *perm, double *xh, int n, hipStream_t s, float *xf) {
  hipLaunchKernelGGL(gather_perm_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, x, perm, xh, xf, n);
}

template <int KMAX>
static void launch_vtx_k(const VTask *t, int nt, const void *vdata, int mode, double *w, double *zpart, hipStream_t s,
                         int waves, const float *xf, const int *active, const VItem *items, int nitems) {
  if (items) {
    if (nitems <= 0) return;
    if (mode == kStoreF32)
      hipLaunchKernelGGL((lowrank_vtx_items_kernel<float, false, KMAX>), dim3(nblk(nitems, waves)), dim3(64 * waves),
                         0, s, t, items, nitems, (const float *)vdata, w, xf, zpart, active);
    else if (mode == kStoreF32Acc64)
      hipLaunchKernelGGL((lowrank_vtx_items_kernel<float, true, KMAX>), dim3(nblk(nitems, waves)), dim3(64 * waves),
                         0, s, t, items, nitems, (const float *)vdata, w, xf, zpart, active);
    else if (mode == kStoreF64NT)
      hipLaunchKernelGGL((lowrank_vtx_items_nt_kernel<KMAX>), dim3(nblk(nitems, waves)), dim3(64 * waves), 0, s, t,
                         items, nitems, (const double *)vdata, w, xf, zpart, active);
    else
      hipLaunchKernelGGL((lowrank_vtx_items_kernel<double, true, KMAX>), dim3(nblk(nitems, waves)), dim3(64 * waves),
                         0, s, t, items, nitems, (const double *)vdata, w, xf, zpart, active);
    return;
  }
  if (mode == kStoreF32)
    hipLaunchKernelGGL((lowrank_vtx_kernel<float, false, KMAX>), dim3(nblk(nt, waves)), dim3(64 * waves), 0, s, t, nt,
                       (const float *)vdata, w, xf, zpart, active);
  else if (mode == kStoreF32Acc64)
    hipLaunchKernelGGL((lowrank_vtx_kernel<float, true, KMAX>), dim3(nblk(nt, waves)), dim3(64 * waves), 0, s, t, nt,
                       (const float *)vdata, w, xf, zpart, active);
  else if (mode == kStoreF64NT)
    hipLaunchKernelGGL((lowrank_vtx_nt_kernel<KMAX>), dim3(nblk(nt, waves)), dim3(64 * waves), 0, s, t, nt,
                       (const double *)vdata, w, xf, zpart, active);
  else
    hipLaunchKernelGGL((lowrank_vtx_kernel<double, true, KMAX>), dim3(nblk(nt, waves)), dim3(64 * waves), 0, s, t,
                       nt, (const double *)vdata, w, xf, zpart, active);
}

int vtx_rank_ceiling(int kmax) { return kmax <= 16 ? 16 : kmax <= 20 ? 20 : kMaxRank; }

void launch_vtx(const VTask *t, int nt, const void *vdata, int mode, double *w, double *zpart, hipStream_t s,
                int waves, const float *xf, const int *active, int kmax, const VItem *items, int nitems) {
  if (nt <= 0) return;
  if (waves != 2 && waves != 8) waves = 4;
  const int kc = vtx_rank_ceiling(kmax);
  if (kc == 16) launch_vtx_k<16>(t, nt, vdata, mode, w, zpart, s, waves, xf, active, items, nitems);
  else if (kc == 20) launch_vtx_k<20>(t, nt, vdata, mode, w, zpart, s, waves, xf, active, items, nitems);
  else launch_vtx_k<kMaxRank>(t, nt, vdata, mode, w, zpart, s, waves, xf, active, items, nitems);
}

template <int R, int KMAX>
static void launch_vtx_near_k(int64_t nb, const VTask *vt, int nvt, const float *vdata, const PTask *pt, int npt,
                              const double *pdata, const int *widx, double *w, double *zpart, double *ypart,
                              hipStream_t s, const int *active) {
  hipLaunchKernelGGL((vtx_near_kernel<R, KMAX>), dim3((unsigned)nb), dim3(256), 0, s, vt, nvt, vdata, pt, npt, pdata,
                     widx, w, zpart, ypart, active);
}

void launch_vtx_near(const VTask *vt, int nvt, const float *vdata, const PTask *pt, int npt, const double *pdata,
                     const int *widx, double *w, double *zpart, double *ypart, hipStream_t s, int rows,
                     const int *active, int kmax) {
  const int64_t nb = (int64_t)nblk(nvt, 4) + npt;
  if (nb <= 0) return;
  const int kc = vtx_rank_ceiling(kmax);
#define DFN_NEAR(R, K) launch_vtx_near_k<R, K>(nb, vt, nvt, vdata, pt, npt, pdata, widx, w, zpart, ypart, s, active)
  if (rows == 4) {
    if (kc == 16) DFN_NEAR(4, 16);
    else if (kc == 20) DFN_NEAR(4, 20);
    else DFN_NEAR(4, kMaxRank);
  } else {
    if (kc == 16) DFN_NEAR(2, 16);
    else if (kc == 20) DFN_NEAR(2, 20);
    else DFN_NEAR(2, kMaxRank);
  }
#undef DFN_NEAR
}

void launch_zreduce(const ZRed *zr, int nb, const double *zpart, double *w, hipStream_t s) {
  if (nb <= 0) return;
  hipLaunchKernelGGL(zreduce_kernel, dim3(nblk((int64_t)nb * 32, 256)), dim3(256), 0, s, zr, nb, zpart, w);
}

void launch_panel(const PTask *t, int nt, const void *pdata, int mode, const int *widx, const float *cscale,
                  const double *w, double *ypart, hipStream_t s, int rows, const int *active) {
  if (nt <= 0) return;
  const bool fp32 = mode == kStoreF32;
  if (mode == kStoreF32Acc64 && rows == 4)
    hipLaunchKernelGGL(panel_kernel_f32w<4>, dim3(nt), dim3(256), 0, s, t, (const float *)pdata, widx, cscale, w,
                       ypart, active);
  else if (mode == kStoreF32Acc64)
    hipLaunchKernelGGL(panel_kernel_f32w<2>, dim3(nt), dim3(256), 0, s, t, (const float *)pdata, widx, cscale, w,
                       ypart, active);
  else if (fp32 && rows == 4)
    hipLaunchKernelGGL(panel_kernel_f32<4>, dim3(nt), dim3(256), 0, s, t, (const float *)pdata, widx, cscale, w, ypart,
                       active);
  else if (fp32)
    hipLaunchKernelGGL(panel_kernel_f32<2>, dim3(nt), dim3(256), 0, s, t, (const float *)pdata, widx, cscale, w, ypart,
                       active);
  else if (rows == 4)
    hipLaunchKernelGGL(panel_kernel_f64<4>, dim3(nt), dim3(256), 0, s, t, (const double *)pdata, widx, w, ypart,
                       active);
  else
    hipLaunchKernelGGL(panel_kernel_f64<2>, dim3(nt), dim3(256), 0, s, t, (const double *)pdata, widx, w, ypart,
                       active);
}

void launch_panel_split(const PTask *t, int nt, const float *hi, const signed char *lo, const int *widx,
                        const float *cscale, const double *w, double *ypart, hipStream_t s, int rows,
                        const int *active) {
  if (nt <= 0) return;
  if (rows == 4)
    hipLaunchKernelGGL(panel_kernel_split<4>, dim3(nt), dim3(256), 0, s, t, hi, lo, widx, cscale, w, ypart, active);
  else
    hipLaunchKernelGGL(panel_kernel_split<2>, dim3(nt), dim3(256), 0, s, t, hi, lo, widx, cscale, w, ypart, active);
}

void launch_finalize(const int *fin_first, const int *fin_idx, const int *perm, const double *ypart, int n, int nops,
                     double scale, double *yT, double *yN, hipStream_t s) {
  hipLaunchKernelGGL(finalize_kernel, dim3(nblk((int64_t)n * nops, 256)), dim3(256), 0, s, fin_first, fin_idx, perm,
                     ypart, n, nops, scale, yT, yN);
}

void launch_mrhs(const MrhsLayout &d, const double *X, int nrhs, int R, double *W, double *zpart, double *ypart,
                 double *Yh, double scale, double *yT, double *yN, hipStream_t s, hipEvent_t *ev) {
  const int n = d.n, groups = R / 64;
  hipLaunchKernelGGL(mrhs_gather_kernel, dim3(nblk(n, 32), R / 32), dim3(256), 0, s, X, d.perm, n, nrhs, R, W);
  if (d.nvt > 0)
    hipLaunchKernelGGL(mrhs_vtx_kernel, dim3(nblk((int64_t)d.nvt * groups, 4)), dim3(256), 0, s, d.vt, d.nvt,
                       d.vdata, W, zpart, R, groups);
  if (d.nzr > 0)
    hipLaunchKernelGGL(mrhs_zreduce_kernel, dim3(nblk((int64_t)d.nzr * 32 * R, 256)), dim3(256), 0, s, d.zr, d.nzr,
                       zpart, W, R);
  if (ev) (void)hipEventRecord(ev[0], s);
  if (d.npt > 0)
    hipLaunchKernelGGL(mrhs_panel_kernel, dim3(d.npt * groups), dim3(256), 0, s, d.pt, d.pdata, d.widx, W, ypart, R,
                       groups);
  if (ev) (void)hipEventRecord(ev[1], s);
  hipLaunchKernelGGL(mrhs_finalize_kernel, dim3(nblk((int64_t)n * d.nops * R, 256)), dim3(256), 0, s, d.fin_first,
                     d.fin_idx, ypart, n, d.nops, R, Yh);
  for (int o = 0; o < d.nops; o++)
    hipLaunchKernelGGL(mrhs_scatter_kernel, dim3(nblk(n, 32), R / 32), dim3(256), 0, s, Yh + (int64_t)o * n * R,
                       d.perm, n, nrhs, R, scale, o == 0 ? yT : yN);
}

int direct_chunks(int n) { return (n + kDirectCols) / kDirectCols; }

void launch_direct3(const double *node, const double *slip, double *part, int n, double scale, double *sT,
                    hipStream_t s) {
  const int nc = direct_chunks(n);
  hipLaunchKernelGGL(direct3_tile_kernel, dim3(nblk(n, 256), nc), dim3(256), 0, s, node, slip, n, part);
  hipLaunchKernelGGL(direct_reduce_kernel, dim3(nblk(n, 64)), dim3(256), 0, s, part, nc, n, scale, sT);
}

void launch_direct2(const double *node, const double *elem, const double *nrm, const double *slip, double *part, int n,
                    double scale, double *sT, double *sN, hipStream_t s) {
  const int nc = direct_chunks(n);
  double *pT = part, *pN = part + (int64_t)nc * n;
  hipLaunchKernelGGL(direct2_tile_kernel, dim3(nblk(n, 256), nc), dim3(256), 0, s, node, elem, nrm, slip, n, pT, pN);
  hipLaunchKernelGGL(direct_reduce_kernel, dim3(nblk(n, 64)), dim3(256), 0, s, pT, nc, n, scale, sT);
  hipLaunchKernelGGL(direct_reduce_kernel, dim3(nblk(n, 64)), dim3(256), 0, s, pN, nc, n, scale, sN);
}

void launch_stage(int stage, const StageArgs &a, hipStream_t s) {
  const dim3 g(nblk(a.n, 256)), b(256);
  switch (stage) {
    case 0: hipLaunchKernelGGL(stage_kernel<0>, g, b, 0, s, a); break;
    case 1: hipLaunchKernelGGL(stage_kernel<1>, g, b, 0, s, a); break;
    case 2: hipLaunchKernelGGL(stage_kernel<2>, g, b, 0, s, a); break;
    case 3: hipLaunchKernelGGL(stage_kernel<3>, g, b, 0, s, a); break;
    case 4: hipLaunchKernelGGL(stage_kernel<4>, g, b, 0, s, a); break;
    case 5: hipLaunchKernelGGL(stage_kernel<5>, g, b, 0, s, a); break;
    default: hipLaunchKernelGGL(stage_kernel<6>, g, b, 0, s, a); break;
  }
}

void launch_ens_stage(int stage, const EnsStageArgs &a, hipStream_t s) {
  const dim3 g(nblk(a.a.n, 256)), b(256);
  switch (stage) {
    case 1: hipLaunchKernelGGL(ens_stage_kernel<1>, g, b, 0, s, a); break;
    case 2: hipLaunchKernelGGL(ens_stage_kernel<2>, g, b, 0, s, a); break;
    case 3: hipLaunchKernelGGL(ens_stage_kernel<3>, g, b, 0, s, a); break;
    case 4: hipLaunchKernelGGL(ens_stage_kernel<4>, g, b, 0, s, a); break;
    case 5: hipLaunchKernelGGL(ens_stage_kernel<5>, g, b, 0, s, a); break;
    default: hipLaunchKernelGGL(ens_stage_kernel<6>, g, b, 0, s, a); break;
  }
}

void launch_ens_restart_stage2(const EnsStageArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(ens_restart_stage2_kernel, dim3(nblk(a.a.n, 256)), dim3(256), 0, s, a);
}

void launch_ens_vstats(const double *V, const uint8_t *fake, const CgNet *nets, int nnet, double *out, hipStream_t s) {
  hipLaunchKernelGGL(ens_vstats_kernel, dim3(nnet), dim3(1024), 0, s, V, fake, nets, out);
}

// One save of the solver output (save_data, src/solver.f90:211-222): column col
// of six (n x ld) row-major fields in buf: V, log state, P, the last shear stress
// rate, sigma_N and the permeability of every element.
__global__ void __launch_bounds__(256) save_output_kernel(const double *Y, const double *P, const double *sT,
                                                          const double *perm, double *buf, int n, int64_t ld,
                                                          int64_t col) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int64_t o = (int64_t)e * ld + col, fs = (int64_t)n * ld;
  buf[o] = Y[e];
  buf[fs + o] = Y[n + e];
  buf[2 * fs + o] = P[e];
  buf[3 * fs + o] = sT[e];
  buf[4 * fs + o] = Y[2 * n + e];
  buf[5 * fs + o] = perm[e];
}

void launch_save_output(const double *Y, const double *P, const double *sT, const double *perm, double *buf, int n,
                        int64_t ld, int64_t col, hipStream_t s) {
  hipLaunchKernelGGL(save_output_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, Y, P, sT, perm, buf, n, ld, col);
}

void launch_monitor(const double *Vold, const double *Vnew, double h, const int *fault_of, int n, double *slip,
                    unsigned long long *fvmax, int nfault, hipStream_t s) {
  hipLaunchKernelGGL(fill_u64_kernel, dim3(nblk(nfault, 256)), dim3(256), 0, s, fvmax, nfault);
  hipLaunchKernelGGL(monitor_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, Vold, Vnew, h, fault_of, n, slip, fvmax);
}

void launch_ens_clear_status(StatusWord *st, const int *active, int nnet, hipStream_t s) {
  hipLaunchKernelGGL(ens_clear_status_kernel, dim3(nblk(nnet, 256)), dim3(256), 0, s, st, active, nnet);
}

void launch_ens_cg_dt(CgSys *sys, int nnet, int nslot, const double *t_try, const double *h, int first_slot,
                      hipStream_t s) {
  hipLaunchKernelGGL(ens_cg_dt_kernel, dim3(nblk((int64_t)nnet * nslot, 256)), dim3(256), 0, s, sys, nnet, nslot,
                     t_try, h, first_slot);
}

void launch_restart_stage2(const double *y, const double *k1, double h, int n3, double *yt, hipStream_t s) {
  hipLaunchKernelGGL(restart_stage2_kernel, dim3(nblk(n3, 256)), dim3(256), 0, s, y, k1, h, n3, yt);
}

void launch_cg(const CgArgs &a, int nsys, hipStream_t s) {
  hipLaunchKernelGGL(cg_kernel, dim3(nsys), dim3(1024), 0, s, a);
}

__global__ void __launch_bounds__(256) stream_copy_kernel(const double2 *__restrict__ a, double2 *__restrict__ b,
                                                          int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

void launch_stream_copy(const void *a, void *b, int64_t bytes, hipStream_t s) {
  const int64_t n = bytes / 16;
  hipLaunchKernelGGL(stream_copy_kernel, dim3(4096), dim3(256), 0, s, (const double2 *)a, (double2 *)b, n);
}

size_t cg_coop_lds_bytes(int cap) {
  return sizeof(double) * (kCgSpan + 2) + (size_t)(cap > 1 ? cap : 1) * (8 * sizeof(double) + 3 * sizeof(int));
}

int cg_coop_capacity(int cap) {
  if (cap > kCgCoopMaxEntries) return 0;
  int per_cu = 0, dev = 0, cus = 0, coop = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, cg_coop_kernel, 256, cg_coop_lds_bytes(cap)) !=
          hipSuccess)
    return 0;
  return per_cu * cus;
}

bool launch_cg_coop(const CgArgs &a, const CrossRows &xr, int nsys, int cap, int max_blocks, unsigned seq,
                    unsigned long long *records, hipStream_t s) {
  if (cap > kCgCoopMaxEntries) return false;
  const size_t sh = cg_coop_lds_bytes(cap);
  int wps = (a.n + kCgSpan - 1) / kCgSpan;
  if (wps > max_blocks) return false;
  // as many systems per cooperative launch as can be resident together
  const int per_launch = max_blocks / wps;
  for (int s0 = 0; s0 < nsys; s0 += per_launch) {
    const int ns = nsys - s0 < per_launch ? nsys - s0 : per_launch;
    CgArgs A = a;
    A.sys = a.sys + s0;
    CrossRows X = xr;
    unsigned long long *rec = records + (size_t)kCgRecordWords * s0 * wps;
    void *args[] = {&A, &X, &wps, &cap, &seq, &rec};
    if (hipLaunchCooperativeKernel((const void *)cg_coop_kernel, dim3(ns * wps), dim3(256), args, sh, s) !=
        hipSuccess)
      return false;
  }
  return true;
}

static bool cg_lds_dispatch(const CgArgs &a, const CrossRows &xr, int nsys, int max_n, hipStream_t s) {
  const size_t sh = sizeof(double) * (size_t)max_n;
  if (max_n <= kCgLdsThreads * 2)
    hipLaunchKernelGGL(cg_lds_kernel<2>, dim3(nsys), dim3(kCgLdsThreads), sh, s, a, xr);
  else if (max_n <= kCgLdsThreads * 4)
    hipLaunchKernelGGL(cg_lds_kernel<4>, dim3(nsys), dim3(kCgLdsThreads), sh, s, a, xr);
  else if (max_n <= kCgLdsThreads * 8)
    hipLaunchKernelGGL(cg_lds_kernel<8>, dim3(nsys), dim3(kCgLdsThreads), sh, s, a, xr);
  else
    return false;
  return true;
}

bool launch_cg_lds(const CgArgs &a, const CrossRows &xr, int nsys, hipStream_t s) {
  return cg_lds_dispatch(a, xr, nsys, a.n, s);
}

// every system's network has at most max_n elements; the kernel is picked by max_n
bool launch_cg_coop_map(const CgArgs &a, const CrossRows &xr, int nwg, int cap, unsigned seq,
                        unsigned long long *records, hipStream_t s) {
  if (cap > kCgCoopMaxEntries || !a.wg_map || nwg < 1) return false;
  const size_t sh = cg_coop_lds_bytes(cap);
  CgArgs A = a;
  CrossRows X = xr;
  int wps = 0;  // taken from the map
  void *args[] = {&A, &X, &wps, &cap, &seq, &records};
  return hipLaunchCooperativeKernel((const void *)cg_coop_kernel, dim3(nwg), dim3(256), args, sh, s) == hipSuccess;
}

void launch_cg_lds_multi(const CgArgs &a, const CrossRows &xr, int nsys, int max_n, hipStream_t s) {
  (void)cg_lds_dispatch(a, xr, nsys, max_n, s);
}

void launch_perm_update(const PermArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(perm_update_kernel, dim3(nblk(a.n, 256)), dim3(256), 0, s, a);
  hipLaunchKernelGGL(perm_x_kernel, dim3(nblk(a.n, 256)), dim3(256), 0, s, a.perm, a.perm_x, a.n);
}

}  // namespace dfn
